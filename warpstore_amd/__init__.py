"""warpstore_amd: MI355X (gfx950) analytic operators behind the engine's GPU operator surface.

The product is the native library libwarpstore_amd.so (C ABI: include/warpstore_amd.h). This
module loads it in-tree and mirrors its entry points with ctypes for tests and the benchmark; it
fails loudly when the library is missing (there is no CPU fallback behind it).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .abi import AggPlan, ColView, DenseSpec, EnvBuilder, ExprEnv, Prog  # noqa: F401

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libwarpstore_amd.so")
DATAGEN_PATH = os.path.join(PKG, "libws_datagen.so")

WS_OK, WS_ERR_OOM, WS_ERR_INTERNAL, WS_ERR_NEEDS_PLAN = 0, 1041, 1105, 1235


class WsError(RuntimeError):
    """A failed C ABI call: .code is the engine's SqlError code (1041, 1105, 1235)."""

    def __init__(self, msg: str, code: int) -> None:
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_gen = None

P, I, I64, U64, SZ, D = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_size_t, C.c_double
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    "ws_last_error": (C.c_char_p, []),
    "ws_init": (I, []), "ws_sm_count": (I, []), "ws_arch": (I, []), "ws_describe": (I, [C.c_char_p, SZ]),
    "ws_dev_alloc": (I, [SZ, PP]), "ws_dev_free": (None, [P]), "ws_tmp_alloc": (I, [SZ, PP]),
    "ws_trim_pool": (None, []), "ws_tmp_free": (None, [P]),
    "ws_pinned_alloc": (I, [SZ, PP]), "ws_pinned_free": (None, [P]),
    "ws_h2d": (I, [P, P, SZ]), "ws_d2h": (I, [P, P, SZ]), "ws_d2h_async": (I, [P, P, SZ]), "ws_d2d": (I, [P, P, SZ]), "ws_sync": (I, []),
    "ws_set_vram_limit": (None, [I64]), "ws_vram_used": (I64, []), "ws_vram_fill": (D, []),
    "ws_place_column": (I, [P, SZ, I, C.POINTER(I), PP]), "ws_release_column": (None, [P, P, SZ, I]),
    "ws_vram_acquire": (P, [P, SZ]), "ws_vram_release": (None, [P, SZ]), "ws_host_map": (P, [P, SZ]),
    "ws_host_unmap": (None, [P]), "ws_vram_limit": (I64, []), "ws_vram_free": (I64, []),
    "ws_patch_column": (P, [P, SZ, I, P, P, I64]), "ws_update_column": (P, [P, P, I64, I, I, P, I64, P]),
    "ws_agg_create": (I, [C.POINTER(AggPlan), I64, C.POINTER(DenseSpec), PP]), "ws_agg_destroy": (None, [P]),
    "ws_agg_add_batch": (I, [P, C.POINTER(ExprEnv), I, I64, P]), "ws_agg_run": (I, [P, C.POINTER(ExprEnv), P]),
    "ws_agg_run_window": (I, [P, C.POINTER(ExprEnv), P]), "ws_agg_batches": (I, [P]), "ws_agg_windows": (I, [P]),
    "ws_agg_rows_in": (I64, [P]), "ws_agg_dense": (I, [P]), "ws_agg_overflowed": (I, [P, C.POINTER(I)]),
    "ws_agg_fetch_begin": (I, [P, C.POINTER(I64)]), "ws_agg_fetch_copy": (I, [P, P, P]),
    "ws_agg_sum_bits_slots": (I, [P, P]), "ws_agg_sum_high": (I, [P, P]),
    "ws_agg_compact": (I, [P, PP, PP, C.POINTER(I64)]), "ws_agg_capacity": (I64, [P]),
    "ws_agg_group_count": (I, [P, C.POINTER(I64)]),
    "ws_agg_set_strategy": (I, [P, I, I64]), "ws_agg_partitioned": (I, [P]),
    "ws_tune_set": (I, [C.c_char_p, I64]), "ws_tune_get": (I, [C.c_char_p, C.POINTER(I64)]),
    "ws_column_stats": (I, [P, I64, P, P]), "ws_col_stats_merge": (None, [P, P]),
    "ws_hll_estimate": (C.c_double, [P]),
    "ws_sizer_create": (I, [C.c_uint64, PP]), "ws_sizer_destroy": (None, [P]), "ws_sizer_get_params": (I, [P, P]),
    "ws_sizer_set_params": (I, [P, P]), "ws_sizer_predict": (I, [P, P, I64, P]),
    "ws_sizer_loss_grad": (I, [P, P, P, I64, P, P]),
    "ws_sizer_train": (I, [P, P, P, I64, I, I, C.c_float, C.c_uint64, P]),
    "ws_sizer_groups": (I, [P, P, C.POINTER(I64)]), "ws_sizing_features": (I, [P, I, P, I64, C.c_double, P]),
    "ws_agg_set_features": (I, [P, P]), "ws_agg_run_batch": (I, [I, P, P, P]),
    "ws_agg_batch_prepare": (I, [I, P, P, P, PP]), "ws_agg_batch_replay": (I, [P, P, C.POINTER(I64), I64, P, P]),
    "ws_agg_batch_launch": (I, [P]), "ws_agg_batch_finish": (I, [P, P, C.POINTER(I64), I64, P, P]),
    "ws_agg_batch_release": (None, [P]),
    "ws_agg_fetch_many": (I, [I, P, P, C.POINTER(I64), I64, P, P]), "ws_sizing_log_size": (I64, []), "ws_sizing_log_copy": (I64, [P, P, I64]),
    "ws_sizing_log_clear": (None, []),
    "ws_select": (I, [C.POINTER(ExprEnv), I, Prog, I64, P, P, P, C.POINTER(I64)]),
    "ws_select_noalloc": (I, [C.POINTER(ExprEnv), I, Prog, I64, P, P, C.POINTER(I64)]),
    "ws_select_multi": (I, [C.POINTER(ExprEnv), I, Prog, I, P, P, P, I, PP, P, P, C.POINTER(I64)]),
    "ws_gather_multi": (I, [P, P, I, I, I, P, I64, P]), "ws_remap_codes": (I, [P, I64, P, I64, P]),
    "ws_bits_to_double": (I, [P, I64, P]), "ws_eval_at_int": (I, [C.POINTER(ExprEnv), Prog, P, I64, P]),
    "ws_distinct_ints": (I, [C.POINTER(ExprEnv), Prog, P, I64, I64, I64, P, I64, C.POINTER(I64)]),
    "ws_bits_from_column": (I, [P, I, I64, I64, I64, C.POINTER(I), PP]),
    "ws_temp_column": (I, [P, I, P, I, I64, C.POINTER(abi.TempColSpec), C.POINTER(D), C.POINTER(D), PP]),
    "ws_eval_at": (I, [C.POINTER(ExprEnv), Prog, P, I64, P]), "ws_gather": (I, [P, I, P, I64, P]),
    "ws_gather_enc": (I, [C.POINTER(ColView), I, P, I64, P]), "ws_key_bitmap": (I, [ColView, I64, I64, I64, P]),
    "ws_bitmap_count": (I, [P, I64, C.POINTER(I64)]),
    "ws_key_bitmap_where": (I, [C.POINTER(ExprEnv), I, Prog, Prog, I, P, P, P, I, I64, I64, P, P, P]),
    "ws_set_bits": (I, [P, P, I64]), "ws_gather_null": (I, [P, I, P, I64, P]),
    "ws_gather_bytes": (I, [P, I, P, I64, P, I]),
    "ws_hash_join": (I, [ColView, I64, ColView, I64, PP, PP, C.POINTER(I64)]),
    "ws_hash_join_parts": (I, [ColView, P, I, ColView, P, I, PP, PP, P, C.POINTER(I64)]),
    "ws_hash_join_outer": (I, [ColView, I64, ColView, I64, PP, PP, I, C.POINTER(I64)]),
    "ws_cross_join": (I, [I64, I64, PP, PP, C.POINTER(I64)]),
    "ws_hash_join_sized": (I, [ColView, I64, ColView, I64, I64, PP, PP, C.POINTER(I64)]),
    "ws_join_features": (I, [P, I64, I64, P, I64, P]), "ws_join_log_add": (None, [P, I64]),
    "ws_join_log_size": (I64, []), "ws_join_log_copy": (I64, [P, P, I64]), "ws_join_log_clear": (None, []),
    "ws_sort_perm": (I, [P, P, I, I64, PP]), "ws_topk_perm": (I, [P, P, I, I64, I64, PP]), "ws_map_rank": (I, [P, P, I64]),
    "ws_sort_perm_host": (I, [P, I, I64, P]),
    "ws_jit_enabled": (I, []), "ws_jit_mode": (I, []), "ws_jit_set_mode": (None, [I]),
    "ws_jit_set_default_mode": (None, [I]), "ws_jit_set_cache_dir": (None, [C.c_char_p]),
    "ws_jit_get": (I, [C.POINTER(ExprEnv), I, C.POINTER(AggPlan), I, C.c_char_p, PP]),
    "ws_jit_program_id": (I, [P]), "ws_jit_launch": (I, [P, I, C.c_uint, C.c_uint, SZ, PP]),
    "ws_prog_is_int": (I, [C.POINTER(ExprEnv), Prog, C.POINTER(I)]), "ws_jit_begin_query": (None, [C.c_char_p]),
    "ws_jit_end_query": (None, [D, C.POINTER(abi.JitQueryInfo)]),
    "ws_jit_plan_cache": (I, [C.POINTER(abi.PlanCacheRow), I]), "ws_jit_source_of": (I, [I, C.c_char_p, SZ]),
    "ws_jit_clear_cache": (None, []),
    "ws_jit_compile_check": (I, [C.POINTER(ExprEnv), I, C.POINTER(AggPlan), I, C.c_char_p, C.c_char_p, SZ,
                                 C.POINTER(SZ)]),
    "ws_prof_enable": (None, [I]), "ws_prof_only": (None, [C.c_char_p]), "ws_prof_reset": (None, []), "ws_prof_collect": (I, [C.POINTER(abi.ProfRow), I]),
    "ws_copy_kernel": (I, [P, P, SZ]),
    "ws_encoded_bytes_max": (SZ, [I64]), "ws_encode_column": (I, [P, I, I64, P, SZ, C.POINTER(I), C.POINTER(SZ)]),
}


def lib():
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with warpstore_amd/build.py "
                               "(there is no CPU fallback for the GPU operators)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def datagen():
    global _gen
    if _gen is None:
        if not os.path.exists(DATAGEN_PATH):
            raise RuntimeError(f"{DATAGEN_PATH} is missing: build it with warpstore_amd/build.py")
        G = C.CDLL(DATAGEN_PATH)
        G.wsg_c1.argtypes = [U64, I64, P, P, P]
        G.wsg_c2.argtypes = [U64, I64, I64, I64, P, P, P, P, P]
        G.wsg_c3_dim.argtypes = [U64, I64, P, P]
        G.wsg_c3_fact.argtypes = [U64, I64, I64, P, P, P, P]
        G.wsg_c4_shapes.argtypes = [U64, I64, P, P, P]
        G.wsg_c4_data.argtypes = [U64, I64, I64, I64, P, P]
        for f in ("wsg_c1", "wsg_c2", "wsg_c3_dim", "wsg_c3_fact", "wsg_c4_shapes", "wsg_c4_data"):
            getattr(G, f).restype = None
        _gen = G
    return _gen


def check(rc: int) -> None:
    if rc != WS_OK:
        raise WsError(lib().ws_last_error().decode(errors="replace"), rc)


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class DevBuf:
    """Device memory from the library's allocator (persistent ws_dev_alloc, or pool ws_tmp_alloc)."""

    def __init__(self, nbytes: int, tmp: bool = False, adopt: int | None = None) -> None:
        self.nbytes = int(nbytes)
        self.tmp = tmp
        if adopt is not None:
            self.p = adopt
            return
        out = C.c_void_p()
        check((lib().ws_tmp_alloc if tmp else lib().ws_dev_alloc)(max(self.nbytes, 1), C.byref(out)))
        self.p = out.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().ws_h2d(b.p, ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        if n:
            check(lib().ws_d2h(ptr(out), self.p, n * dtype.itemsize))
        return out

    def free(self) -> None:
        if self.p:
            (lib().ws_tmp_free if self.tmp else lib().ws_dev_free)(self.p)
            self.p = None

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:
            pass


def device_count_ok() -> bool:
    """True when the library can initialize a GPU here."""
    try:
        return lib().ws_init() == WS_OK
    except Exception:
        return False


def prof_rows() -> dict:
    rows = (abi.ProfRow * 64)()
    n = lib().ws_prof_collect(rows, 64)
    return {rows[i].name.decode(): (rows[i].launches, rows[i].total_ms) for i in range(min(n, 64))}
