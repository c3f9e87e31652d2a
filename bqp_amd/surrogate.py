"""Learned gravity surrogate over the HIP C ABI (``bqp_surrogate_*``).

No reference counterpart: the reference evaluates every harmonic term
exactly (crates/satpy_core/src/gravity.rs:2638).  The surrogate keeps the
field up to degree ``nlow`` exact (FP64) and learns the rest,

    16 features of the body-fixed position -> 4 x [256, SiLU] -> 3,

trained on the GPU with the library's own FP32 forward/backward/Adam
kernels on seeded residuals computed with the exact FP64 recursion, and
evaluated on MFMA in BF16 or FP8 E4M3 (the three 256 x 256 layers, with
per-layer scales), FP32 accumulation.  In the propagator
(``precision="surrogate"``) every ``check_every``-th evaluation of a row is
compared with the exact field; a difference above ``threshold`` m/s^2 puts
that row on the exact field for the rest of its arc, and positions outside
the radius band always use the exact field.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as nat
from .device import DeviceArray

N_FEATURES, WIDTH, N_OUT, N_LAYERS = 16, 256, 3, 5
LAYER_DIMS = [(16, 256), (256, 256), (256, 256), (256, 256), (256, 3)]  # (in, out)


def param_slices():
    """(W slice, b slice, (out, in)) per layer of the flat FP32 parameter vector."""
    out, off = [], 0
    for i, o in LAYER_DIMS:
        w = slice(off, off + i * o)
        off += i * o
        b = slice(off, off + o)
        off += o
        out.append((w, b, (o, i)))
    return out


N_PARAMS = param_slices()[-1][1].stop  # 202,499
MLP_FLOPS = 2.0 * (16 * 256 + 3 * 256 * 256 + 256 * 3)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _vp(x):
    return C.c_void_p(x.data_ptr()) if x is not None else None


class Surrogate:
    """The surrogate of one Propagator (its context, field and stream)."""

    def __init__(self, prop, nlow: int = 4, r_lo: float = 6578.0e3, r_hi: float = 8378.0e3, seed: int = 0):
        self.prop = prop
        self._lib = prop._lib
        self.nlow, self.r_lo, self.r_hi = int(nlow), float(r_lo), float(r_hi)
        n = int(self._lib.bqp_surrogate_param_count())
        if n != N_PARAMS:
            raise nat.BqpError(f"library reports {n} parameters, expected {N_PARAMS}")
        prop._check(self._lib.bqp_surrogate_init(prop._ctx, self.nlow, self.r_lo, self.r_hi, int(seed)))
        prop.surrogate = self

    # -------------------------------------------------------- parameters
    @property
    def out_scale(self) -> np.ndarray:
        s = np.empty(3)
        self.prop._check(self._lib.bqp_surrogate_get_params(self.prop._ctx, None,
                                                            s.ctypes.data_as(C.POINTER(C.c_double))))
        return s

    def params(self) -> np.ndarray:
        p = np.empty(N_PARAMS, dtype=np.float32)
        self.prop._check(self._lib.bqp_surrogate_get_params(self.prop._ctx, _fp(p), None))
        return p

    def set_params(self, params=None, out_scale=None):
        pp = None if params is None else np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if pp is not None and pp.size != N_PARAMS:
            raise ValueError(f"need {N_PARAMS} parameters")
        sp = None if out_scale is None else np.ascontiguousarray(out_scale, dtype=np.float64).reshape(3)
        self.prop._check(self._lib.bqp_surrogate_set_params(
            self.prop._ctx, _fp(pp) if pp is not None else None,
            sp.ctypes.data_as(C.POINTER(C.c_double)) if sp is not None else None))

    def save(self, path: str):
        np.savez(path, params=self.params(), out_scale=self.out_scale,
                 meta=np.array([self.nlow, self.r_lo, self.r_hi]))

    def load(self, path: str):
        with np.load(path, allow_pickle=False) as z:
            meta = z["meta"]
            if int(meta[0]) != self.nlow or float(meta[1]) != self.r_lo or float(meta[2]) != self.r_hi:
                raise ValueError("saved surrogate was set up with another nlow / radius band")
            self.set_params(z["params"], z["out_scale"])

    # ------------------------------------------------------------- data
    def dataset(self, n: int, seed: int = 1, positions: bool = False):
        """Seeded device-resident training rows: (features (n,16) f32,
        scaled targets (n,3) f32, positions (n,3) f64 or None)."""
        feat = DeviceArray((n, N_FEATURES), np.float32)
        targ = DeviceArray((n, N_OUT), np.float32)
        pos = DeviceArray((n, 3), np.float64) if positions else None
        self.prop._check(self._lib.bqp_surrogate_dataset_dev(self.prop._ctx, int(n), int(seed), _vp(pos),
                                                             _vp(feat), _vp(targ)))
        return feat, targ, pos

    # --------------------------------------------------------- training
    def train(self, feat, targ, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        """`steps` Adam steps at a constant learning rate; returns the
        per-step minibatch losses."""
        loss = np.empty(max(int(steps), 1), dtype=np.float32)
        self.prop._check(self._lib.bqp_surrogate_train_dev(self.prop._ctx, _vp(feat), _vp(targ),
                                                           int(feat.shape[0]), int(steps), int(batch), float(lr),
                                                           int(seed), _fp(loss)))
        return loss[:steps]

    def fit(self, feat, targ, steps: int, batch: int = 4096, lr_max: float = 2e-3, lr_min: float = 2e-5,
            chunk: int = 250, seed: int = 0) -> np.ndarray:
        """Cosine-decayed learning rate, applied in constant chunks."""
        out, done, k = [], 0, 0
        while done < steps:
            m = min(chunk, steps - done)
            mid = (done + 0.5 * m) / steps
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + math.cos(math.pi * mid))
            out.append(self.train(feat, targ, m, batch, lr, seed=seed * 1000003 + k))
            done += m
            k += 1
        return np.concatenate(out) if out else np.empty(0, np.float32)

    def loss_grad(self, feat, targ):
        """Full-batch mean squared error and its gradient (FP32, device)."""
        loss = C.c_double(0.0)
        g = np.empty(N_PARAMS, dtype=np.float32)
        self.prop._check(self._lib.bqp_surrogate_loss_grad_dev(self.prop._ctx, _vp(feat), _vp(targ),
                                                               int(feat.shape[0]), C.byref(loss), _fp(g)))
        return loss.value, g

    # -------------------------------------------------------- inference
    def pack(self, fmt: str = "bf16", calib=None):
        """Pack the parameters for MFMA inference ('bf16' or 'fp8'; fp8
        calibrates activation scales on the `calib` feature rows)."""
        if fmt not in ("bf16", "fp8"):
            raise ValueError("fmt must be 'bf16' or 'fp8'")
        n = int(calib.shape[0]) if calib is not None else 0
        self.prop._check(self._lib.bqp_surrogate_pack(self.prop._ctx, nat.MLP_FORMATS[fmt], _vp(calib), n))
        self.fmt = fmt

    def infer(self, feat, fmt: str = "bf16", out=None):
        """Network outputs (n,3) in scaled units on the device; returns
        (DeviceArray, kernel_ms)."""
        n = int(feat.shape[0])
        if out is None:
            out = DeviceArray((n, N_OUT), np.float32)
        ms = C.c_double(0.0)
        self.prop._check(self._lib.bqp_surrogate_infer_dev(self.prop._ctx, nat.MLP_FORMATS[fmt], _vp(feat), n,
                                                           _vp(out), C.byref(ms)))
        return out, ms.value

    def set_check(self, check_every: int = 100, threshold: float = 1e-9):
        self.prop._check(self._lib.bqp_surrogate_set_check(self.prop._ctx, int(check_every), float(threshold)))

    def last_stats(self) -> dict:
        s = nat.SurrogateStats()
        self.prop._check(self._lib.bqp_surrogate_last_stats(self.prop._ctx, C.byref(s)))
        return s.as_dict()
