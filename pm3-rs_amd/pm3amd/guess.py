"""Learned initial density (include/pm3_guess.h): the network handle, the device-resident training
set, GPU training, and the kernel-level hooks used by the parity tests.

The network predicts density blocks for every directed atom pair within 4 Angstrom of molecules
made of H, C and O; the SCF keeps the assembled guess only when its first-iteration commutator
norm is not larger than the superposition guess's, so converged answers are those of the plain
SCF.  Training runs entirely on the GPU (hand-written bf16 matrix-core kernels, Adam).
"""

import ctypes
import os
from typing import Optional

import numpy as np

from . import Batch, Molecule, Pm3Parameters, _dp, _raise, native

NF, NH, NO = 96, 128, 16
NPARAM = NF * NH + NH * NH + NH + NH * NO + NO
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "guess_hco.bin")


class _CTrainOptions(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("batch_molecules", ctypes.c_int32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("lr_final", ctypes.c_float), ("decay_steps", ctypes.c_int32)]


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u16(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


class GuessModel:
    """Network of the learned guess (fp32 master weights on the device, bf16 fragments)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def create(cls, seed=1):
        h = ctypes.c_void_p()
        _raise(native().pm3_guess_model_create(int(seed), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        _raise(native().pm3_guess_model_load(os.fsencode(path), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def default(cls):
        """The network trained by this engine and shipped with the package."""
        if not os.path.exists(DEFAULT_MODEL):
            raise FileNotFoundError(f"no trained guess model at {DEFAULT_MODEL}")
        return cls.load(DEFAULT_MODEL)

    def save(self, path):
        _raise(native().pm3_guess_model_save(self.handle, os.fsencode(path)))

    def params(self):
        out = np.zeros(NPARAM, dtype=np.float32)
        _raise(native().pm3_guess_model_get_params(self.handle, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAM
        _raise(native().pm3_guess_model_set_params(self.handle, _fp(p)))

    @property
    def steps(self):
        return native().pm3_guess_model_steps(self.handle)

    def mlp(self, x, dy=None):
        """Network outputs y (n x 16) of bf16 inputs x (n x 96, uint16 bits); with dy also the
        parameter gradient of sum(dy * y)."""
        x = np.ascontiguousarray(x, dtype=np.uint16)
        n = len(x)
        y = np.zeros((n, NO), dtype=np.float32)
        if dy is None:
            _raise(native().pm3_guess_mlp(self.handle, n, _u16(x), _fp(y), None, None))
            return y
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        _raise(native().pm3_guess_mlp(self.handle, n, _u16(x), _fp(y), _fp(dy), _fp(g)))
        return y, g

    def density(self, molecule: Molecule, params: Optional[Pm3Parameters] = None):
        """(learned guess density n x n, |[F,P]|^2 learned, |[F,P]|^2 superposition)."""
        params = params or Pm3Parameters.standard()
        cm = molecule._c()
        n = native().pm3_query_nao(ctypes.byref(cm), params._h)
        P = np.zeros((n, n))
        cl, cs = ctypes.c_double(), ctypes.c_double()
        _raise(native().pm3_guess_density(self.handle, ctypes.byref(cm), params._h, _dp(P), ctypes.byref(cl),
                                          ctypes.byref(cs)))
        return P, cl.value, cs.value

    def close(self):
        if getattr(self, "handle", None):
            native().pm3_guess_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def features(molecule: Molecule, params: Optional[Pm3Parameters] = None):
    """(src, dst, x): directed pairs (caller atom indices) and their bf16 inputs (uint16 bits)."""
    params = params or Pm3Parameters.standard()
    cm = molecule._c()
    n = ctypes.c_int32(0)
    ip = ctypes.POINTER(ctypes.c_int32)
    _raise(native().pm3_guess_features(ctypes.byref(cm), params._h, ctypes.byref(n), None, None, None))
    src = np.zeros(max(n.value, 1), dtype=np.int32)
    dst = np.zeros(max(n.value, 1), dtype=np.int32)
    x = np.zeros((max(n.value, 1), NF), dtype=np.uint16)
    _raise(native().pm3_guess_features(ctypes.byref(cm), params._h, ctypes.byref(n), src.ctypes.data_as(ip),
                                       dst.ctypes.data_as(ip), _u16(x)))
    k = n.value
    return src[:k], dst[:k], x[:k]


class GuessData:
    """Device-resident training set built from solved batches (converged densities)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _raise(native().pm3_guess_data_create(ctypes.byref(h)))
        self.handle = h.value

    def add(self, batch: Batch):
        """Appends the converged H/C/O molecules of a solved batch (created with want_density)."""
        n = ctypes.c_int64(0)
        _raise(native().pm3_guess_data_add(self.handle, batch._h, ctypes.byref(n)))
        return n.value

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _raise(native().pm3_guess_data_stats(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"molecules": a.value, "atoms": b.value, "pairs": c.value}

    def close(self):
        if getattr(self, "handle", None):
            native().pm3_guess_data_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train(model: GuessModel, data: GuessData, steps, batch_molecules=1024, lr=1.0e-3, seed=1, beta1=0.9,
          beta2=0.999, eps=1.0e-8, lr_final=1.0e-5, decay_steps=0):
    """Adam steps on the GPU; returns (loss before each step, seconds).  decay_steps > 0: cosine
    step size from lr to lr_final over the model's first decay_steps optimiser steps."""
    o = _CTrainOptions(int(steps), int(batch_molecules), float(lr), float(beta1), float(beta2), float(eps), int(seed),
                       float(lr_final), int(decay_steps))
    losses = np.zeros(int(steps), dtype=np.float32)
    sec = ctypes.c_double()
    _raise(native().pm3_guess_train(model.handle, data.handle, ctypes.cast(ctypes.pointer(o), ctypes.c_void_p),
                                    _fp(losses), ctypes.byref(sec)))
    return losses, sec.value


def gradient(model: GuessModel, data: GuessData):
    """(parameter gradient, loss) of the whole set as one step, no update."""
    g = np.zeros(NPARAM, dtype=np.float32)
    loss = ctypes.c_double()
    _raise(native().pm3_guess_grad(model.handle, data.handle, _fp(g), ctypes.byref(loss)))
    return g, loss.value


def evaluate(model: GuessModel, data: GuessData):
    """Per-atom loss of the model over the whole set."""
    out = ctypes.c_double()
    _raise(native().pm3_guess_eval(model.handle, data.handle, ctypes.byref(out)))
    return out.value
