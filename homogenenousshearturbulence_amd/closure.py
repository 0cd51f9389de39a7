"""Python binding of the learned subgrid closure (include/hst_closure.h).

The closure has no counterpart in the reference.  A pointwise network
(9 filtered velocity gradients -> 6 subgrid stresses, 9-64-64-6 tanh) is
trained on the GPU by the library's own forward/backward kernels on the
matrix cores (bfloat16 operands, FP32 accumulation) and added to the
products of a coarse run.  There is no CPU fallback: every call goes through
libhstgpu.so and raises HstError on failure.

    fine = hst.Solver(dns_deck, hst.MIXED)          # the DNS
    c = Closure(coarse_deck, capacity=1 << 22)
    fine.nonlinear(); c.filter(fine.time); c.sample(fine.time)   # per snapshot
    c.normalise(); c.init_params(1); c.train(epochs, 65536, 1e-3)
    coarse = hst.Solver(coarse_deck, hst.FP64); coarse.set_V(c.field_velocity())
    c.attach()                                        # coarse.timestep adds tau
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import hst
from .deck import Deck

NPARAM = 5190
NORM = 30
U, PBAR, TAU = 0, 3, 9
STRESS_ORDER = ("uu", "vv", "ww", "uv", "vw", "uw")

EXPORTS = (
    "hst_closure_filter_weights", "hst_closure_create", "hst_closure_free", "hst_closure_filter",
    "hst_closure_get_field", "hst_closure_set_field", "hst_closure_sample", "hst_closure_set_dataset",
    "hst_closure_get_dataset", "hst_closure_count", "hst_closure_normalise", "hst_closure_set_norm",
    "hst_closure_get_norm", "hst_closure_init_params", "hst_closure_set_params", "hst_closure_get_params",
    "hst_closure_grad", "hst_closure_train", "hst_closure_predict", "hst_closure_stress", "hst_closure_attach",
)

# parameter blocks of the flat array: W1[64][9] b1[64] W2[64][64] b2[64] W3[6][64] b3[6]
BLOCKS = (("W1", (64, 9)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (6, 64)), ("b3", (6,)))

_bound = False


def lib() -> ctypes.CDLL:
    global _bound
    L = hst.lib()
    if _bound:
        return L
    vp, d, i, i64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64
    u64 = ctypes.c_uint64
    sig = {
        "hst_closure_filter_weights": [i, i, vp, vp],
        "hst_closure_create": [vp, i64, ctypes.POINTER(vp)],
        "hst_closure_free": [vp],
        "hst_closure_filter": [vp, d],
        "hst_closure_get_field": [vp, i, vp],
        "hst_closure_set_field": [vp, i, vp],
        "hst_closure_sample": [vp, d, ctypes.POINTER(i64)],
        "hst_closure_set_dataset": [vp, i64, vp, vp],
        "hst_closure_get_dataset": [vp, i64, i64, vp, vp],
        "hst_closure_count": [vp, ctypes.POINTER(i64)],
        "hst_closure_normalise": [vp, vp],
        "hst_closure_set_norm": [vp, vp],
        "hst_closure_get_norm": [vp, vp],
        "hst_closure_init_params": [vp, u64],
        "hst_closure_set_params": [vp, vp],
        "hst_closure_get_params": [vp, vp],
        "hst_closure_grad": [vp, i64, i64, vp, ctypes.POINTER(d)],
        "hst_closure_train": [vp, i, i, d, u64, i64, vp, vp],
        "hst_closure_predict": [vp, i64, vp, vp],
        "hst_closure_stress": [vp, d],
        "hst_closure_attach": [vp, i],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _bound = True
    return L


def filter_weights(ny_fine: int, ny_coarse: int):
    """first_row[k], weights[k, 0..7] of the y top-hat (host only)."""
    j0 = np.zeros(ny_coarse, dtype=np.int32)
    w = np.zeros((ny_coarse, 8))
    hst._check(lib().hst_closure_filter_weights(ny_fine, ny_coarse, j0.ctypes.data, w.ctypes.data),
               "hst_closure_filter_weights")
    return j0, w


def unpack(params: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in BLOCKS:
        n = int(np.prod(shape))
        out[name] = params[o:o + n].reshape(shape)
        o += n
    return out


class Closure:
    """A closure context on the coarse grid of `deck`."""

    def __init__(self, deck: Deck, capacity: int = 0):
        self.deck = deck
        self._c = deck.to_c()
        h = ctypes.c_void_p()
        hst._check(lib().hst_closure_create(ctypes.byref(self._c), int(capacity), ctypes.byref(h)),
                   "hst_closure_create")
        self.h = h
        self.shape = (deck.nx + 1, 2 * deck.nz + 1, deck.ny + 4)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            lib().hst_closure_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str) -> None:
        hst._check(rc, what)

    # ---------------------------------------------------------------- fields
    def filter(self, time: float) -> None:
        """Filter the live state's velocity and products (after a product pass)."""
        self._chk(lib().hst_closure_filter(self.h, float(time)), "hst_closure_filter")

    def field(self, which: int) -> np.ndarray:
        a = np.zeros(self.shape, dtype=np.complex128)
        self._chk(lib().hst_closure_get_field(self.h, which, a.ctypes.data), "hst_closure_get_field")
        return a

    def set_field(self, which: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.shape
        self._chk(lib().hst_closure_set_field(self.h, which, a.ctypes.data), "hst_closure_set_field")

    def velocity(self) -> np.ndarray:
        return np.stack([self.field(U + i) for i in range(3)])

    def tau(self) -> np.ndarray:
        return np.stack([self.field(TAU + p) for p in range(6)])

    # ---------------------------------------------------------------- data set
    def sample(self, time: float) -> int:
        n = ctypes.c_int64(0)
        self._chk(lib().hst_closure_sample(self.h, float(time), ctypes.byref(n)), "hst_closure_sample")
        return n.value

    def count(self) -> int:
        n = ctypes.c_int64(0)
        self._chk(lib().hst_closure_count(self.h, ctypes.byref(n)), "hst_closure_count")
        return n.value

    def set_dataset(self, x: np.ndarray, y: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        assert x.shape[1:] == (9,) and y.shape == (x.shape[0], 6)
        self._chk(lib().hst_closure_set_dataset(self.h, x.shape[0], x.ctypes.data, y.ctypes.data),
                  "hst_closure_set_dataset")

    def dataset(self, first: int = 0, n: int | None = None):
        if n is None:
            n = self.count() - first
        x = np.zeros((n, 9), dtype=np.float32)
        y = np.zeros((n, 6), dtype=np.float32)
        self._chk(lib().hst_closure_get_dataset(self.h, first, n, x.ctypes.data, y.ctypes.data),
                  "hst_closure_get_dataset")
        return x, y

    # ---------------------------------------------------------------- network
    def normalise(self) -> np.ndarray:
        nm = np.zeros(NORM)
        self._chk(lib().hst_closure_normalise(self.h, nm.ctypes.data), "hst_closure_normalise")
        return nm

    def set_norm(self, nm: np.ndarray) -> None:
        nm = np.ascontiguousarray(nm, dtype=np.float64)
        self._chk(lib().hst_closure_set_norm(self.h, nm.ctypes.data), "hst_closure_set_norm")

    def norm(self) -> np.ndarray:
        nm = np.zeros(NORM)
        self._chk(lib().hst_closure_get_norm(self.h, nm.ctypes.data), "hst_closure_get_norm")
        return nm

    def init_params(self, seed: int) -> None:
        self._chk(lib().hst_closure_init_params(self.h, int(seed)), "hst_closure_init_params")

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (NPARAM,)
        self._chk(lib().hst_closure_set_params(self.h, p.ctypes.data), "hst_closure_set_params")

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, dtype=np.float32)
        self._chk(lib().hst_closure_get_params(self.h, p.ctypes.data), "hst_closure_get_params")
        return p

    def grad(self, first: int, n: int):
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        self._chk(lib().hst_closure_grad(self.h, first, n, g.ctypes.data, ctypes.byref(loss)), "hst_closure_grad")
        return g, loss.value

    def train(self, epochs: int, batch: int, lr: float, seed: int = 1, n_train: int = 0):
        loss = np.zeros(epochs)
        val = np.zeros(epochs)
        self._chk(lib().hst_closure_train(self.h, epochs, batch, float(lr), int(seed), int(n_train),
                                          loss.ctypes.data, val.ctypes.data), "hst_closure_train")
        return loss, val

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 6), dtype=np.float32)
        self._chk(lib().hst_closure_predict(self.h, x.shape[0], x.ctypes.data, y.ctypes.data), "hst_closure_predict")
        return y

    def stress(self, time: float) -> np.ndarray:
        self._chk(lib().hst_closure_stress(self.h, float(time)), "hst_closure_stress")
        return self.tau()

    def attach(self, on: bool = True) -> None:
        self._chk(lib().hst_closure_attach(self.h, 1 if on else 0), "hst_closure_attach")
