"""Value net standing in for river chance nodes in depth-limited turn solves
(the reference's cfvnet BoundaryNet, cfvnet/src/model/network.rs:10-90 and
python/cfvnet/model.py:1-50), on the device through the psr_mlp_* C ABI:
BF16 MFMA forward / backward, masked Huber loss, AdamW.

Encoding (one row per (ranges, 5-card board, pot, stack) situation):
  input  2 x 1326 range weights (each range scaled to mean 1 over its
         non-zero combos), 52 board one-hot, pot / (pot + stack),
         stack / (pot + stack)                                 -> 2706
  output 2 x 1326 pot-relative values: for hand i of player p, its expected
         payoff against the opponent range in units of half the pot
         (the value a depth-boundary leaf multiplies by half_pot /
         num_combinations and the unblocked opponent reach)   -> 2652
"""
import ctypes as C

import numpy as np

from .range_solver import PsrError, _check, load_library

NUM_COMBOS = 1326
INPUT_DIM = 2 * NUM_COMBOS + 52 + 2
OUTPUT_DIM = 2 * NUM_COMBOS
DEFAULT_DIMS = (INPUT_DIM, 512, 512, 512, 512, OUTPUT_DIM)


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ValueNet:
    """An MLP of the given layer widths (GELU hidden layers, linear output)."""

    def __init__(self, dims=DEFAULT_DIMS, max_batch=1024, seed=0, device=0):
        self.dims = tuple(int(d) for d in dims)
        self.max_batch = int(max_batch)
        arr = (C.c_int * len(self.dims))(*self.dims)
        h = C.c_void_p()
        _check(load_library().psr_mlp_create(device, arr, len(self.dims), self.max_batch, seed, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            load_library().psr_mlp_destroy(self._h)
            self._h = None

    def num_params(self) -> int:
        return _check(load_library().psr_mlp_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        _check(load_library().psr_mlp_get_params(self._h, _f(out), len(out)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(load_library().psr_mlp_set_params(self._h, _f(p), len(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != self.dims[0]:
            raise PsrError(f"input must be n x {self.dims[0]}")
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        _check(load_library().psr_mlp_forward(self._h, _f(x), x.shape[0], _f(y)))
        return y

    def gradients(self, x, y, mask, delta=1.0):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_double()
        _check(load_library().psr_mlp_gradients(self._h, _f(x), _f(y), _f(m), x.shape[0], delta, C.byref(loss), _f(g),
                                                len(g)))
        return loss.value, g

    def set_dataset(self, x, y, mask):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        _check(load_library().psr_mlp_set_dataset(self._h, _f(x), _f(y), _f(m), x.shape[0]))

    def train(self, epochs, batch, lr=1e-3, weight_decay=1e-4, delta=1.0, seed=0) -> np.ndarray:
        out = np.zeros(epochs, np.float64)
        _check(load_library().psr_mlp_train(self._h, epochs, batch, lr, weight_decay, delta, seed,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_grad_clip(self, max_norm: float):
        """Clip every training step's gradient to L2 norm max_norm over all
        parameters (torch.nn.utils.clip_grad_norm_; the reference trainer
        uses 1.0, cfvnet/src/config.rs:284-285).  0 turns it off."""
        _check(load_library().psr_mlp_set_grad_clip(self._h, float(max_norm)))

    def last_grad_norm(self) -> float:
        """The gradient norm of the last clipped training step."""
        v = C.c_float()
        _check(load_library().psr_mlp_last_grad_norm(self._h, C.byref(v)))
        return v.value

    def last_ms(self) -> float:
        return float(load_library().psr_mlp_last_ms(self._h))
