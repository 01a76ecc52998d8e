"""The learned face flux from Python (over the pmf_surrogate_* C ABI).

A 16 -> 64 -> 64 -> 64 -> 9 MLP (tanh GELU; bfloat16 operands on MFMA, float32
accumulation and master weights) that predicts the HLLD dissipation relative to the
central flux on faces the jump detector calls smooth. Training data are face-state /
flux pairs dumped by the plain solver; training runs on the GPU (Adam).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import MHDError, check, lib

THR = 0.02  # gate: relative jumps of rho, p and |B| below 2%


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Surrogate:
    def __init__(self, state, seed: int = 0):
        if state.dtype != np.float32:
            raise MHDError("the flux surrogate runs on float32 states")
        self.state = state
        h = ctypes.c_void_p()
        check(lib().pmf_surrogate_create(state.handle, int(seed), ctypes.byref(h)))
        self._h = h
        self.nparams = lib().pmf_surrogate_nparams()
        self.count = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pmf_surrogate_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def weights(self) -> np.ndarray:
        w = np.empty(self.nparams, dtype=np.float32)
        check(lib().pmf_surrogate_get_weights(self._h, _p(w), w.size))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(lib().pmf_surrogate_set_weights(self._h, _p(w), w.size))

    def reserve(self, capacity: int):
        check(lib().pmf_surrogate_reserve(self._h, int(capacity)))
        self.count = 0

    def dump(self, state, dt, ch, thr=THR, keep=1.0, seed=0) -> int:
        """Append the gated faces of the state's current fields (kept with probability keep)."""
        n = ctypes.c_int64()
        check(lib().pmf_surrogate_dump(self._h, state.handle, float(dt), float(ch), float(thr), float(keep),
                                       int(seed), ctypes.byref(n)))
        self.count = n.value
        return n.value

    def dataset(self, first=0, count=None):
        count = self.count - first if count is None else count
        X = np.empty((count, 16), dtype=np.float32)
        Y = np.empty((count, 9), dtype=np.float32)
        S = np.empty((count, 18), dtype=np.float32)
        check(lib().pmf_surrogate_dataset(self._h, int(first), int(count), _p(X), _p(Y), _p(S)))
        return X, Y, S

    def upload(self, X, Y, S=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        Sp = None if S is None else _p(np.ascontiguousarray(S, dtype=np.float32))
        check(lib().pmf_surrogate_upload(self._h, _p(X), _p(Y), Sp, X.shape[0]))
        self.count = X.shape[0]

    def grad(self, first=0, count=None, bf16=True):
        count = self.count - first if count is None else count
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        check(lib().pmf_surrogate_grad(self._h, int(first), int(count), 1 if bf16 else 0, _p(g), ctypes.byref(loss)))
        return loss.value, g

    def predict(self, first=0, count=None, bf16=True):
        count = self.count - first if count is None else count
        Y = np.empty((count, 9), dtype=np.float32)
        check(lib().pmf_surrogate_predict(self._h, int(first), int(count), 1 if bf16 else 0, _p(Y)))
        return Y

    def train(self, steps, batch=4096, lr=1e-3, seed=0, bf16=True):
        losses = np.empty(max(int(steps), 1), dtype=np.float64)
        check(lib().pmf_surrogate_train(self._h, int(batch), int(steps), float(lr), int(seed), 1 if bf16 else 0,
                                        _p(losses)))
        return losses[:steps]

    def face_flux(self, mode, first=0, count=None, thr=THR):
        """Face fluxes of dataset rows: mode 'hlld', 'surrogate' or 'central'."""
        m = {"hlld": 0, "surrogate": 1, "central": 2}[mode]
        count = self.count - first if count is None else count
        F = np.empty((count, 9), dtype=np.float32)
        check(lib().pmf_surrogate_face_flux(self._h, int(first), int(count), m, float(thr), _p(F)))
        return F


def step_surrogate(state, net: Surrogate, dt, ch, glm_cr=0.18, thr=THR, count=False):
    """step! with the learned flux on gated faces (fused kernel). Returns the number of
    network-solved faces when count=True (this waits for the GPU)."""
    n = ctypes.c_int64()
    check(lib().pmf_step_surrogate(state.handle, net.handle, float(dt), float(ch), float(glm_cr), float(thr),
                                   ctypes.byref(n) if count else None))
    return n.value if count else None
