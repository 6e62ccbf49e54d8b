"""BF16 MFMA multilayer perceptron (the library's sk_mlp_*): exact-erf GELU
hidden layers, FP32 master weights, mean-squared-error training with Adam.
Everything runs in the HIP library; inputs and outputs are device tensors."""

import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._device import ptr, require_gpu, stream


def _destroy(h):
    try:
        _lib.load().sk_mlp_destroy(h)
    except Exception:
        pass


class MLP:
    def __init__(self, dims, max_batch=1024, seed=0):
        require_gpu()
        self.dims = [int(d) for d in dims]
        self.max_batch = int(max_batch)
        arr = (ctypes.c_int * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        _lib.call("sk_mlp_create", len(self.dims) - 1, arr, self.max_batch, ctypes.byref(h))
        self._h = h
        self._fin = weakref.finalize(self, _destroy, h.value)
        self.n_params = int(_lib.load().sk_mlp_num_params(h))
        self.set_params(self.init_params(seed))

    def init_params(self, seed):
        """W, b ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) per layer (seeded), in the flat layout."""
        rng = np.random.default_rng(seed)
        out = []
        for k, n in zip(self.dims[:-1], self.dims[1:]):
            s = 1.0 / np.sqrt(k)
            out.append(rng.uniform(-s, s, n * k))
            out.append(rng.uniform(-s, s, n))
        return np.concatenate(out).astype(np.float32)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _lib.call("sk_mlp_set_params", self._h, p.ctypes.data_as(ctypes.c_void_p), stream())

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        _lib.call("sk_mlp_get_params", self._h, p.ctypes.data_as(ctypes.c_void_p), stream())
        return p

    def get_grad(self):
        g = np.empty(self.n_params, np.float32)
        _lib.call("sk_mlp_get_grad", self._h, g.ctypes.data_as(ctypes.c_void_p), stream())
        return g

    def forward(self, X, out=None):
        X = X.contiguous()
        B = X.shape[0]
        Y = torch.empty((B, self.dims[-1]), dtype=torch.float32, device=X.device) if out is None else out
        _lib.call("sk_mlp_forward", self._h, ptr(X), B, ptr(Y), stream())
        return Y

    def backward(self, X, T, want_loss=True):
        loss = ctypes.c_double()
        _lib.call("sk_mlp_backward", self._h, ptr(X.contiguous()), ptr(T.contiguous()), X.shape[0],
                  ctypes.byref(loss) if want_loss else None, stream())
        return loss.value if want_loss else None

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _lib.call("sk_mlp_adam", self._h, float(lr), float(beta1), float(beta2), float(eps), stream())

    def fit(self, X, T, epochs=10, batch=256, lr=1e-3, seed=0, lr_decay=1.0, log=None):
        """Mini-batch Adam on device tensors X (n, d0), T (n, dL); returns per-epoch mean losses."""
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(X.device)
            tot, cnt = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                last = s + 2 * batch > n
                loss = self.backward(X[idx], T[idx], want_loss=last)
                self.adam_step(lr)
                if last:
                    tot += loss
                    cnt += 1
            losses.append(tot / max(cnt, 1))
            if log:
                log(ep, losses[-1])
            lr *= lr_decay
        return losses
