"""Training the learned sizing models from the logs the drivers and aggregations keep.

Two logs, one model each (the same MLP, ws_sizer, 16 -> 32 -> 32 -> 1, trained on the GPU):
  - groups: (ws_sizing_features, exact group count) of every aggregation that carried features;
  - joins:  (ws_join_features, exact output pairs) of every hash join the drivers ran.
A trained model is installed in the drivers with queries.set_sizers(); the plain rules stay the
fallback (an under-estimate retries bigger or writes again), so results never depend on it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import check, lib

N_IN = 16


class Sizer:
    """A ws_sizer handle (freed with close())."""

    def __init__(self, seed: int = 9) -> None:
        self.h = C.c_void_p()
        check(lib().ws_sizer_create(seed, C.byref(self.h)))
        self.loss = None
        self.samples = 0

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 400, batch: int = 32, lr: float = 3e-3,
              seed: int = 9) -> "Sizer":
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().ws_sizer_train(self.h, x.ctypes.data, t.ctypes.data, len(t), epochs, min(batch, max(len(t), 1)), lr, seed,
                                   loss.ctypes.data))
        self.loss = float(loss[0])
        self.samples = len(t)
        return self

    def predict(self, x: np.ndarray) -> np.ndarray:
        """Predicted counts (groups or pairs) of feature rows x (n x 16)."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        y = np.zeros(len(x), np.float32)
        check(lib().ws_sizer_predict(self.h, x.ctypes.data, len(x), y.ctypes.data))
        return np.maximum(1.0, np.exp2(32.0 * y.astype(np.float64)) - 1)

    def close(self) -> None:
        if self.h:
            lib().ws_sizer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def log_copy(kind: str):
    """(features n x 16, targets log2(count + 1) / 32) of the groups or joins log."""
    L = lib()
    n = L.ws_sizing_log_size() if kind == "groups" else L.ws_join_log_size()
    x = np.zeros((max(n, 1), N_IN), np.float32)
    t = np.zeros(max(n, 1), np.float32)
    copy = L.ws_sizing_log_copy if kind == "groups" else L.ws_join_log_copy
    n = copy(x.ctypes.data, t.ctypes.data, n)
    return x[:n], t[:n]


def log_clear(kind: str) -> None:
    (lib().ws_sizing_log_clear if kind == "groups" else lib().ws_join_log_clear)()


def train_from_log(kind: str, epochs: int = 400, seed: int = 9) -> Sizer:
    x, t = log_copy(kind)
    if len(t) == 0:
        raise RuntimeError(f"the {kind} log is empty: run queries with features first")
    return Sizer(seed).train(x, t, epochs=epochs, seed=seed)
