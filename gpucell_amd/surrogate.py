"""Python handle on the learned membrane substepper's network
(gc_Mlp* in libgpucell_hip.so, include/gpucell/surrogate.h).

Device buffers are torch CUDA tensors (plumbing only: every computation is
a library kernel).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

IN, OUT = 36, 24


class Mlp:
    def __init__(self, seed: int = 1):
        self.lib = abi.load()
        st = C.c_int(0)
        self.h = self.lib.gc_MlpCreate(C.c_uint64(seed), C.byref(st))
        if not self.h:
            raise abi.GcError(f"gc_MlpCreate failed ({st.value}): {self.lib.gc_GetLastError().decode()}")
        self.nparam = self.lib.gc_MlpParamCount()

    def close(self):
        if getattr(self, "h", None):
            self.lib.gc_MlpDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        abi.check(self.lib.gc_MlpGetParams(self.h, p.ctypes.data, self.nparam), "MlpGetParams")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        abi.check(self.lib.gc_MlpSetParams(self.h, p.ctypes.data, self.nparam), "MlpSetParams")

    def forward(self, x):
        """x: CUDA float32 tensor [B][36] -> [B][24]."""
        import torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], OUT), dtype=torch.float32, device=x.device)
        abi.check(self.lib.gc_MlpForward(self.h, x.data_ptr(), y.data_ptr(), int(x.shape[0])),
                  "MlpForward")
        return y

    def gradient(self, x, dy):
        import torch
        g = torch.empty(self.nparam, dtype=torch.float32, device=x.device)
        abi.check(self.lib.gc_MlpGradient(self.h, x.contiguous().data_ptr(), dy.contiguous().data_ptr(),
                                          int(x.shape[0]), g.data_ptr()), "MlpGradient")
        return g

    def train(self, x, y, epochs: int, batch: int = 4096, lr: float = 2e-3, seed: int = 7):
        """Adam on MSE; returns the per-epoch mean training loss."""
        loss = np.zeros(max(epochs, 1), np.float32)
        abi.check(self.lib.gc_MlpTrain(self.h, x.contiguous().data_ptr(), y.contiguous().data_ptr(),
                                       int(x.shape[0]), int(batch), int(epochs), C.c_float(lr),
                                       C.c_uint64(seed), loss.ctypes.data), "MlpTrain")
        return loss[:epochs]
