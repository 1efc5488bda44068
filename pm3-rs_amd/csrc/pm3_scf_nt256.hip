// The batched SCF kernel (pm3_engine.hip) compiled for 256-thread workgroups, 2 resident per CU
// at the same 256-VGPR budget, for molecules whose LDS needs fit 80 KiB.  Only the kernel
// and its entry points (pm3variant::nt256_*) come from this translation unit.
#define PM3_NT 256
#define PM3_KERNEL_VARIANT nt256
#define PM3_SCF_KERNEL scf_batch_kernel_256
#include "pm3_engine.hip"
