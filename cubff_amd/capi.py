"""ctypes binding of the C-ABI in include/cubff_hip.h (libcubff_hip.so).

Plumbing for tests, the smoke check and bench.py.  The library is loaded from
cubff_amd/lib/; if it is missing, import fails loudly: there is no CPU
fallback for any of these entry points.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcubff_hip.so")

PAIR_SHUFFLE, PAIR_FIXED, PAIR_CYCLIC, PAIR_2D, PAIR_HOST = 0, 1, 2, 3, 4
FLOW_OFF, FLOW_ON, FLOW_AUTO, FLOW_DEFER = 0, 1, 2, 3  # cubff_set_pipeline modes

# Every symbol include/cubff_hip.h declares.
EXPORTS = [
    "cubff_last_error", "cubff_lang_id", "cubff_lang_name", "cubff_num_langs",
    "cubff_ctx_create", "cubff_ctx_destroy", "cubff_init_programs",
    "cubff_write_soup", "cubff_read_soup", "cubff_set_shuffle",
    "cubff_set_interactions", "cubff_pair", "cubff_run_epoch",
    "cubff_run_epochs", "cubff_run_epochs_timed", "cubff_set_pipeline", "cubff_flow_stats", "cubff_set_defer_budget",
    "cubff_set_resolver_idle", "cubff_set_selfrep_chunk",
    "cubff_byte_histogram",
    "cubff_read_counters", "cubff_reset_counters",
    "cubff_read_pair_ops", "cubff_read_shuffle", "cubff_selfrep",
    "cubff_eval_tapes", "cubff_eval_tapes_noalloc", "cubff_eval_selfrep", "cubff_trace_bff", "cubff_sync",
    "cubff_trace_steps", "cubff_eval_selfrep_trace",
    "cubff_splitmix64", "cubff_epoch_seed",
    "cubff_batch_last_error", "cubff_batch_create", "cubff_batch_destroy",
    "cubff_batch_init", "cubff_batch_set_launch", "cubff_batch_get_launch",
    "cubff_batch_run_epochs", "cubff_batch_run_epochs_timed",
    "cubff_batch_autotune", "cubff_batch_sync", "cubff_batch_read_soup",
    "cubff_batch_write_soup", "cubff_batch_read_counters",
    "cubff_batch_reset_counters", "cubff_batch_read_pair_stats",
    "cubff_batch_read_shuffle", "cubff_batch_sched_stats",
    "cubff_batch_set_pipeline", "cubff_batch_set_defer_budget", "cubff_batch_defer_stats",
    "cubff_batch_next_epoch",
    "cubff_mlp_last_error", "cubff_mlp_num_params", "cubff_mlp_create",
    "cubff_mlp_destroy", "cubff_mlp_set_params", "cubff_mlp_get_params",
    "cubff_mlp_forward", "cubff_mlp_predict", "cubff_mlp_predict_pieces", "cubff_mlp_gradients",
    "cubff_mlp_train",
]

SCHED_PLAIN, SCHED_LEARNED = 0, 1

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i = ctypes.c_int

_lib = None


class CubffError(RuntimeError):
    pass


def load():
    """Loads libcubff_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CubffError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "cubff_last_error": (ctypes.c_char_p, []),
        "cubff_lang_id": (_i, [ctypes.c_char_p]),
        "cubff_lang_name": (ctypes.c_char_p, [_i]),
        "cubff_num_langs": (_i, []),
        "cubff_ctx_create": (_i, [ctypes.POINTER(_vp), _i, _u64, _i]),
        "cubff_ctx_destroy": (_i, [_vp]),
        "cubff_init_programs": (_i, [_vp, _u64, _i, _u64]),
        "cubff_write_soup": (_i, [_vp, _u64, _u8p, _u64]),
        "cubff_read_soup": (_i, [_vp, _u64, _u8p, _u64]),
        "cubff_set_shuffle": (_i, [_vp, _u32p, _u64]),
        "cubff_set_interactions": (_i, [_vp, _u64p, _u32p]),
        "cubff_pair": (_i, [_vp, _i, _u64, _u64]),
        "cubff_run_epoch": (_i, [_vp, _u64, ctypes.c_uint32]),
        "cubff_run_epochs": (_i, [_vp, _i, _u64, _u64, _u64, ctypes.c_uint32]),
        "cubff_run_epochs_timed": (_i, [_vp, _i, _u64, _u64, _u64, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]),
        "cubff_set_pipeline": (_i, [_vp, _i, ctypes.c_double]),
        "cubff_flow_stats": (_i, [_vp, _u64p, _u64p, _u64p]),
        "cubff_set_defer_budget": (_i, [_vp, ctypes.c_uint32]),
        "cubff_set_selfrep_chunk": (_i, [_vp, ctypes.c_uint64]),
        "cubff_set_resolver_idle": (_i, [_vp, ctypes.c_double]),
        "cubff_byte_histogram": (_i, [_vp, _u64p]),
        "cubff_read_counters": (_i, [_vp, _u64p, _u64p, _u64p]),
        "cubff_reset_counters": (_i, [_vp]),
        "cubff_read_pair_ops": (_i, [_vp, _u32p, _u32p, _u64]),
        "cubff_read_shuffle": (_i, [_vp, _u32p, _u64p]),
        "cubff_selfrep": (_i, [_vp, _u64, _u64, _u64p]),
        "cubff_eval_tapes": (_i, [_i, _i, _u8p, _u64, _u64, _u32p, _u32p]),
        "cubff_eval_tapes_noalloc": (_i, [_i, _i, _u8p, _u64, _u64, _u32p, _u32p]),
        "cubff_trace_bff": (_i, [_i, _i, _u8p, _u64, _u8p, _u64, _u64p, _u64p]),
        "cubff_eval_selfrep": (_i, [_i, _i, _u8p, _u64, _u64, _u64, _u64p]),
        "cubff_eval_selfrep_trace": (_i, [_i, _i, _u8p, _u64, _u64, _u64, _u64p, _u8p]),
        "cubff_trace_steps": (_i, [_i, _i, _u8p, _u64, _u8p, _u64, _u64p, _u64p, _u64p]),
        "cubff_sync": (_i, [_vp]),
        "cubff_splitmix64": (_u64, [_u64]),
        "cubff_epoch_seed": (_u64, [_u64, _u64]),
        "cubff_batch_last_error": (ctypes.c_char_p, []),
        "cubff_batch_create": (_i, [ctypes.POINTER(_vp), ctypes.c_uint32,
                                    ctypes.POINTER(_i), _u64p, _i]),
        "cubff_batch_destroy": (_i, [_vp]),
        "cubff_batch_init": (_i, [_vp, _u64p, _u32p, _i]),
        "cubff_batch_set_launch": (_i, [_vp, _i, ctypes.c_uint32]),
        "cubff_batch_get_launch": (_i, [_vp, ctypes.POINTER(_i), _u32p]),
        "cubff_batch_run_epochs": (_i, [_vp, _u64, _u64, _i, _i]),
        "cubff_batch_run_epochs_timed": (_i, [_vp, _u64, _u64, _i, _i,
                                              ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_float)]),
        "cubff_batch_autotune": (_i, [_vp, _u64, _u64, _i, ctypes.POINTER(_i), _u32p,
                                      ctypes.POINTER(ctypes.c_float)]),
        "cubff_batch_sync": (_i, [_vp]),
        "cubff_batch_read_soup": (_i, [_vp, ctypes.c_uint32, _u8p]),
        "cubff_batch_write_soup": (_i, [_vp, ctypes.c_uint32, _u8p]),
        "cubff_batch_read_counters": (_i, [_vp, ctypes.c_uint32, _u64p, _u64p, _u64p]),
        "cubff_batch_reset_counters": (_i, [_vp]),
        "cubff_batch_read_pair_stats": (_i, [_vp, ctypes.c_uint32, _u32p, _u32p, _u32p]),
        "cubff_batch_read_shuffle": (_i, [_vp, ctypes.c_uint32, _u32p]),
        "cubff_batch_sched_stats": (_i, [_vp, _u64p, ctypes.POINTER(ctypes.c_float), _u32p]),
        "cubff_batch_set_pipeline": (_i, [_vp, _i]),
        "cubff_batch_set_defer_budget": (_i, [_vp, ctypes.c_uint32]),
        "cubff_batch_defer_stats": (_i, [_vp, _u64p]),
        "cubff_batch_next_epoch": (_i, [_vp, _u64p]),
        "cubff_mlp_last_error": (ctypes.c_char_p, []),
        "cubff_mlp_num_params": (_u64, []),
        "cubff_mlp_create": (_i, [ctypes.POINTER(_vp), _u64, _u64, _i]),
        "cubff_mlp_destroy": (_i, [_vp]),
        "cubff_mlp_set_params": (_i, [_vp, _f32p]),
        "cubff_mlp_get_params": (_i, [_vp, _f32p]),
        "cubff_mlp_forward": (_i, [_vp, _u8p, _u64, _f32p]),
        "cubff_mlp_predict": (_i, [_vp, _u8p, _u64, _f32p]),
        "cubff_mlp_predict_pieces": (_i, [_vp, _u8p, _u64, _u32p, _u64, _f32p]),
        "cubff_mlp_gradients": (_i, [_vp, _u8p, _f32p, _u64, _f32p, _f32p]),
        "cubff_mlp_train": (_i, [_vp, _u8p, _f32p, _u64, ctypes.c_uint32, ctypes.c_float,
                                 _f32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise CubffError(load().cubff_last_error().decode())


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def lang_id(name: str) -> int:
    i = load().cubff_lang_id(name.encode())
    if i < 0:
        raise CubffError(f"Unknown language `{name}`")
    return i


def splitmix64(x: int) -> int:
    return int(load().cubff_splitmix64(x & (2**64 - 1)))


def epoch_seed(params_seed: int, x: int) -> int:
    return int(load().cubff_epoch_seed(params_seed & (2**64 - 1), x & (2**64 - 1)))


def eval_tapes(lang: str, tapes: np.ndarray, steps: int = 8192, device: int = 0):
    """Evaluates a (count, 128) uint8 array of tapes on the GPU.
    Returns (final_tapes, ops, steps)."""
    t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128).copy()
    n = t.shape[0]
    ops = np.zeros(n, dtype=np.uint32)
    st = np.zeros(n, dtype=np.uint32)
    _check(load().cubff_eval_tapes(lang_id(lang), device, _ptr(t, _u8p), n, steps,
                                   _ptr(ops, _u32p), _ptr(st, _u32p)))
    return t, ops, st


NOALLOC_MAX_TAPES = 16384


def eval_tapes_noalloc(lang: str, tapes: np.ndarray, steps: int = 8192, device: int = 0):
    """eval_tapes through cubff_eval_tapes_noalloc (no device allocation; at
    most NOALLOC_MAX_TAPES tapes).  Returns (final_tapes, ops, steps)."""
    t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128).copy()
    n = t.shape[0]
    ops = np.zeros(n, dtype=np.uint32)
    st = np.zeros(n, dtype=np.uint32)
    _check(load().cubff_eval_tapes_noalloc(lang_id(lang), device, _ptr(t, _u8p), n, steps,
                                           _ptr(ops, _u32p), _ptr(st, _u32p)))
    return t, ops, st


def eval_selfrep(lang: str, prog: bytes | np.ndarray, epoch: int, seed: int,
                 device: int = 0) -> int:
    p = np.frombuffer(bytes(prog), dtype=np.uint8).copy()
    out = ctypes.c_uint64(0)
    _check(load().cubff_eval_selfrep(lang_id(lang), device, _ptr(p, _u8p), p.size,
                                     epoch, seed, ctypes.byref(out)))
    return int(out.value)


STEP_RECORD_BYTES = 272


def trace_steps(lang: str, tape: bytes | np.ndarray, steps: int, cap: int | None = None,
                device: int = 0):
    """Step records of one Evaluate on the GPU: (records uint8[k, 272], final
    tape, ops, steps run).  Record: u32 pc, r0, r1, r2, tape, stack."""
    t = np.frombuffer(bytes(tape), dtype=np.uint8).copy()
    if t.size != 128:
        raise ValueError("tape must be exactly 128 bytes long")
    cap = steps if cap is None else cap
    rec = np.zeros((max(cap, 1), STEP_RECORD_BYTES), dtype=np.uint8)
    k, ops, run = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(load().cubff_trace_steps(lang_id(lang), device, _ptr(t, _u8p), steps,
                                    _ptr(rec, _u8p), cap, ctypes.byref(k), ctypes.byref(ops),
                                    ctypes.byref(run)))
    return rec[:k.value].copy(), t, int(ops.value), int(run.value)


def eval_selfrep_trace(lang: str, prog: bytes | np.ndarray, epoch: int, seed: int,
                       device: int = 0):
    """(score, tapes uint8[13, 5, 2, 128]) of the single-program check."""
    p = np.frombuffer(bytes(prog), dtype=np.uint8).copy()
    out = ctypes.c_uint64(0)
    tapes = np.zeros((13, 5, 2, 128), dtype=np.uint8)
    _check(load().cubff_eval_selfrep_trace(lang_id(lang), device, _ptr(p, _u8p), p.size,
                                           epoch, seed, ctypes.byref(out), _ptr(tapes, _u8p)))
    return int(out.value), tapes


def trace_bff(lang: str, tape: bytes | np.ndarray, steps: int = 8192, device: int = 0):
    """Step trace of one BFF-family tape on the GPU: (file image bytes, ops)."""
    t = np.frombuffer(bytes(tape), dtype=np.uint8).copy()
    if t.size != 128:
        raise ValueError("tape must be exactly 128 bytes long")
    out = np.zeros(12 + 140 * steps, dtype=np.uint8)
    nbytes, ops = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(load().cubff_trace_bff(lang_id(lang), device, _ptr(t, _u8p), steps, _ptr(out, _u8p),
                                  out.size, ctypes.byref(nbytes), ctypes.byref(ops)))
    return out[:nbytes.value].tobytes(), int(ops.value)


class Soup:
    """One device soup behind a cubff_ctx."""

    def __init__(self, lang: str, num_programs: int, device: int = 0):
        L = load()
        self.lang, self.n = lang, num_programs
        h = _vp()
        _check(L.cubff_ctx_create(ctypes.byref(h), lang_id(lang), num_programs, device))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().cubff_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_programs(self, seed_value: int, zero_init: bool = False, count=None):
        count = self.n if count is None else count
        _check(load().cubff_init_programs(self._h, seed_value, int(zero_init), count))

    def write_soup(self, data: np.ndarray, offset: int = 0):
        d = np.ascontiguousarray(data, dtype=np.uint8)
        _check(load().cubff_write_soup(self._h, offset, _ptr(d, _u8p), d.size))

    def read_soup(self) -> np.ndarray:
        out = np.zeros(self.n * 64, dtype=np.uint8)
        _check(load().cubff_read_soup(self._h, 0, _ptr(out, _u8p), out.size))
        return out

    def set_shuffle(self, s: np.ndarray, num_indices: int | None = None):
        sa = np.ascontiguousarray(s, dtype=np.uint32)
        k = self.n // 2 if num_indices is None else num_indices
        _check(load().cubff_set_shuffle(self._h, _ptr(sa, _u32p), k))

    def set_interactions(self, row_off: np.ndarray, nbr: np.ndarray):
        self._ro = np.ascontiguousarray(row_off, dtype=np.uint64)
        self._nb = np.ascontiguousarray(nbr, dtype=np.uint32)
        if self._nb.size == 0:
            self._nb = np.zeros(1, dtype=np.uint32)
        _check(load().cubff_set_interactions(self._h, _ptr(self._ro, _u64p),
                                             _ptr(self._nb, _u32p)))

    def pair(self, mode: int, params_seed: int, epoch: int):
        _check(load().cubff_pair(self._h, mode, params_seed, epoch))

    def run_epoch(self, seed_value: int, mutation_prob: int):
        _check(load().cubff_run_epoch(self._h, seed_value, mutation_prob))

    def run_epochs(self, mode: int, params_seed: int, first_epoch: int,
                   count: int, mutation_prob: int):
        _check(load().cubff_run_epochs(self._h, mode, params_seed, first_epoch,
                                       count, mutation_prob))

    def set_pipeline(self, mode, wait_limit_s: float = 10.0):
        """Multi-epoch runs: FLOW_ON / True (dataflow kernel), FLOW_OFF /
        False (one launch per epoch) or FLOW_AUTO (the default)."""
        _check(load().cubff_set_pipeline(self._h, int(mode), float(wait_limit_s)))

    def flow_stats(self):
        """(interactions, claims) of the dataflow kernel and interactions
        deferred to the resolver, since creation."""
        a, b, c = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(load().cubff_flow_stats(self._h, ctypes.byref(a), ctypes.byref(b),
                                       ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def byte_histogram(self) -> np.ndarray:
        c = np.zeros(256, dtype=np.uint64)
        _check(load().cubff_byte_histogram(self._h, _ptr(c, _u64p)))
        return c

    def set_defer_budget(self, budget: int):
        _check(load().cubff_set_defer_budget(self._h, budget))

    def set_selfrep_chunk(self, programs: int):
        _check(load().cubff_set_selfrep_chunk(self._h, programs))

    def set_resolver_idle(self, seconds: float):
        _check(load().cubff_set_resolver_idle(self._h, seconds))

    def run_epochs_timed(self, mode: int, params_seed: int, first_epoch: int,
                         count: int, mutation_prob: int):
        """Returns (total_ms, epoch_kernel_ms) from HIP events on the
        context stream."""
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        _check(load().cubff_run_epochs_timed(self._h, mode, params_seed, first_epoch,
                                             count, mutation_prob, ctypes.byref(a),
                                             ctypes.byref(b)))
        return float(a.value), float(b.value)

    def read_counters(self, with_steps: bool = False):
        """(ops, num_runs) since the last reset; (ops, num_runs, steps) with
        with_steps."""
        a, b, c = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(load().cubff_read_counters(self._h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c)))
        if with_steps:
            return int(a.value), int(b.value), int(c.value)
        return int(a.value), int(b.value)

    def reset_counters(self):
        _check(load().cubff_reset_counters(self._h))

    def read_pair_ops(self):
        ops = np.zeros(self.n // 2, dtype=np.uint32)
        st = np.zeros(self.n // 2, dtype=np.uint32)
        _check(load().cubff_read_pair_ops(self._h, _ptr(ops, _u32p), _ptr(st, _u32p),
                                          self.n // 2))
        return ops, st

    def read_shuffle(self):
        s = np.zeros(self.n, dtype=np.uint32)
        k = ctypes.c_uint64(0)
        _check(load().cubff_read_shuffle(self._h, _ptr(s, _u32p), ctypes.byref(k)))
        return s, int(k.value)

    def selfrep(self, seed_value: int, count: int | None = None) -> np.ndarray:
        count = self.n if count is None else count
        out = np.zeros(max(count, 1), dtype=np.uint64)
        _check(load().cubff_selfrep(self._h, seed_value, count, _ptr(out, _u64p)))
        return out[:count]

    def sync(self):
        _check(load().cubff_sync(self._h))


def _bcheck(rc: int):
    if rc != 0:
        raise CubffError(load().cubff_batch_last_error().decode())


class Batch:
    """Many independent soups advanced together (cubff_batch)."""

    def __init__(self, langs, num_programs, device: int = 0):
        L = load()
        self.langs = list(langs)
        self.n = [int(x) for x in num_programs]
        if len(self.langs) != len(self.n):
            raise ValueError("langs and num_programs differ in length")
        ids = (ctypes.c_int * len(self.langs))(*[lang_id(x) for x in self.langs])
        ns = np.array(self.n, dtype=np.uint64)
        h = _vp()
        _bcheck(L.cubff_batch_create(ctypes.byref(h), len(self.n), ids, _ptr(ns, _u64p),
                                     device))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().cubff_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, params_seeds, mutation_probs, zero_init: bool = False):
        sd = np.array([x & (2**64 - 1) for x in params_seeds], dtype=np.uint64)
        pr = np.array(mutation_probs, dtype=np.uint32)
        _bcheck(load().cubff_batch_init(self._h, _ptr(sd, _u64p), _ptr(pr, _u32p),
                                        int(zero_init)))

    def set_launch(self, wg: int, tapes_per_lane: int):
        _bcheck(load().cubff_batch_set_launch(self._h, wg, tapes_per_lane))

    def get_launch(self):
        w, t = ctypes.c_int(0), ctypes.c_uint32(0)
        _bcheck(load().cubff_batch_get_launch(self._h, ctypes.byref(w), ctypes.byref(t)))
        return int(w.value), int(t.value)

    def run_epochs(self, first_epoch: int, count: int, sched: int = SCHED_PLAIN,
                   use_graph: bool = False):
        _bcheck(load().cubff_batch_run_epochs(self._h, first_epoch, count, sched,
                                              int(use_graph)))

    def run_epochs_timed(self, first_epoch: int, count: int, sched: int = SCHED_PLAIN,
                         use_graph: bool = False):
        """(total ms, interpreter ms) of `count` epochs; with use_graph the
        captured epoch is replayed and the interpreter time is -1."""
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        _bcheck(load().cubff_batch_run_epochs_timed(self._h, first_epoch, count, sched,
                                                    int(use_graph), ctypes.byref(a),
                                                    ctypes.byref(b)))
        return float(a.value), float(b.value)

    def autotune(self, first_epoch: int, epochs_per_config: int, sched: int = SCHED_PLAIN):
        w, t = ctypes.c_int(0), ctypes.c_uint32(0)
        tab = (ctypes.c_float * 12)()
        _bcheck(load().cubff_batch_autotune(self._h, first_epoch, epochs_per_config, sched,
                                            ctypes.byref(w), ctypes.byref(t), tab))
        return int(w.value), int(t.value), [float(x) for x in tab]

    def sync(self):
        _bcheck(load().cubff_batch_sync(self._h))

    def read_soup(self, i: int) -> np.ndarray:
        out = np.zeros(self.n[i] * 64, dtype=np.uint8)
        _bcheck(load().cubff_batch_read_soup(self._h, i, _ptr(out, _u8p)))
        return out

    def write_soup(self, i: int, data: np.ndarray):
        d = np.ascontiguousarray(data, dtype=np.uint8)
        assert d.size == self.n[i] * 64
        _bcheck(load().cubff_batch_write_soup(self._h, i, _ptr(d, _u8p)))

    def read_counters(self, i: int):
        a, b, c = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _bcheck(load().cubff_batch_read_counters(self._h, i, ctypes.byref(a),
                                                 ctypes.byref(b), ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def reset_counters(self):
        _bcheck(load().cubff_batch_reset_counters(self._h))

    def read_pair_stats(self, i: int, with_work: bool = False):
        """(ops, steps) of the last epoch's pairs; (ops, steps, work) with
        with_work."""
        ops = np.zeros(self.n[i] // 2, dtype=np.uint32)
        st = np.zeros(self.n[i] // 2, dtype=np.uint32)
        wk = np.zeros(self.n[i] // 2, dtype=np.uint32)
        _bcheck(load().cubff_batch_read_pair_stats(self._h, i, _ptr(ops, _u32p),
                                                   _ptr(st, _u32p), _ptr(wk, _u32p)))
        return (ops, st, wk) if with_work else (ops, st)

    def read_shuffle(self, i: int) -> np.ndarray:
        s = np.zeros(self.n[i], dtype=np.uint32)
        _bcheck(load().cubff_batch_read_shuffle(self._h, i, _ptr(s, _u32p)))
        return s

    def set_pipeline(self, mode: int):
        """FLOW_AUTO / FLOW_DEFER (deferring segments for runs of two or more
        epochs without graph replay) or FLOW_OFF (interpreter rounds only)."""
        _bcheck(load().cubff_batch_set_pipeline(self._h, int(mode)))

    def set_defer_budget(self, budget: int):
        _bcheck(load().cubff_batch_set_defer_budget(self._h, budget))

    def next_epoch(self) -> int:
        """The epoch after the last one run."""
        d = ctypes.c_uint64(0)
        _bcheck(load().cubff_batch_next_epoch(self._h, ctypes.byref(d)))
        return int(d.value)

    def defer_stats(self) -> int:
        """Interactions handed to the resolver since creation."""
        d = ctypes.c_uint64(0)
        _bcheck(load().cubff_batch_defer_stats(self._h, ctypes.byref(d)))
        return int(d.value)

    def sched_stats(self):
        """(trainings, last training loss, slots per length bucket)."""
        t, l = ctypes.c_uint64(0), ctypes.c_float(0)
        c = np.zeros(8, dtype=np.uint32)
        _bcheck(load().cubff_batch_sched_stats(self._h, ctypes.byref(t), ctypes.byref(l),
                                               _ptr(c, _u32p)))
        return int(t.value), float(l.value), c


def _mcheck(rc: int):
    if rc != 0:
        raise CubffError(load().cubff_mlp_last_error().decode())


class Mlp:
    """The scheduler's MLP cost model on its own (cubff_mlp)."""

    def __init__(self, capacity: int, init_seed: int = 0x5eed, device: int = 0):
        h = _vp()
        _mcheck(load().cubff_mlp_create(ctypes.byref(h), capacity, init_seed, device))
        self._h = h
        self.num_params = int(load().cubff_mlp_num_params())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().cubff_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        _mcheck(load().cubff_mlp_get_params(self._h, _ptr(out, _f32p)))
        return out

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, dtype=np.float32)
        _mcheck(load().cubff_mlp_set_params(self._h, _ptr(a, _f32p)))

    def forward(self, tapes: np.ndarray) -> np.ndarray:
        t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128)
        out = np.zeros(t.shape[0], dtype=np.float32)
        _mcheck(load().cubff_mlp_forward(self._h, _ptr(t, _u8p), t.shape[0], _ptr(out, _f32p)))
        return out

    def predict(self, tapes: np.ndarray) -> np.ndarray:
        """The scheduler's fused inference kernel."""
        t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128)
        out = np.zeros(t.shape[0], dtype=np.float32)
        _mcheck(load().cubff_mlp_predict(self._h, _ptr(t, _u8p), t.shape[0], _ptr(out, _f32p)))
        return out

    def predict_pieces(self, pieces: np.ndarray, index: np.ndarray) -> np.ndarray:
        """The same kernel on tapes given as pairs of 64-byte pieces: tape r
        is pieces[index[2r]] followed by pieces[index[2r + 1]]."""
        p = np.ascontiguousarray(pieces, dtype=np.uint8).reshape(-1, 64)
        ix = np.ascontiguousarray(index, dtype=np.uint32).reshape(-1)
        if ix.size % 2:
            raise ValueError("index needs two entries per tape")
        out = np.zeros(ix.size // 2, dtype=np.float32)
        _mcheck(load().cubff_mlp_predict_pieces(self._h, _ptr(p, _u8p), p.shape[0],
                                                _ptr(ix, _u32p), out.size, _ptr(out, _f32p)))
        return out

    def gradients(self, tapes: np.ndarray, targets: np.ndarray):
        t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128)
        y = np.ascontiguousarray(targets, dtype=np.float32)
        loss = ctypes.c_float(0)
        g = np.zeros(self.num_params, dtype=np.float32)
        _mcheck(load().cubff_mlp_gradients(self._h, _ptr(t, _u8p), _ptr(y, _f32p), t.shape[0],
                                           ctypes.byref(loss), _ptr(g, _f32p)))
        return float(loss.value), g

    def train(self, tapes: np.ndarray, targets: np.ndarray, steps: int, lr: float = 1e-3):
        t = np.ascontiguousarray(tapes, dtype=np.uint8).reshape(-1, 128)
        y = np.ascontiguousarray(targets, dtype=np.float32)
        losses = np.zeros(steps, dtype=np.float32)
        _mcheck(load().cubff_mlp_train(self._h, _ptr(t, _u8p), _ptr(y, _f32p), t.shape[0],
                                       steps, lr, _ptr(losses, _f32p)))
        return losses
