// The batched SCF kernel (pm3_engine.hip) compiled for 128-thread workgroups, 4 resident per CU
// at the same 256-VGPR budget, for molecules whose LDS needs fit 40 KiB.  Only the kernel
// and its entry points (pm3variant::nt128_*) come from this translation unit.
#define PM3_NT 128
#define PM3_KERNEL_VARIANT nt128
#define PM3_SCF_KERNEL scf_batch_kernel_128
#include "pm3_engine.hip"
