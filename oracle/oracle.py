"""ctypes view of the CPU oracle (oracle/libcubff_oracle.so).

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker or as the
timed CPU baseline.  The product path (cubff_amd/) never imports it.

Each wrapper restates a reference function; the C source cites file:line.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CUBFF_ORACLE_LIB names another build of the same code, such as the
# AddressSanitizer / UBSan build with the debug checks (make -C oracle san;
# the process then needs the sanitizer runtimes preloaded).
LIB_PATH = os.environ.get("CUBFF_ORACLE_LIB") or os.path.join(HERE, "libcubff_oracle.so")

LANGS = [
    "bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit",
    "bff_perm", "bff_selfmove", "forth", "forthcopy", "forthtrivial",
    "forthtrivial_reset", "subleq", "rsubleq4",
]

_lib = None
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, u32, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        L.orc_splitmix64.restype = u64
        L.orc_splitmix64.argtypes = [u64]
        L.orc_epoch_seed.restype = u64
        L.orc_epoch_seed.argtypes = [u64, u64]
        L.orc_lang_id.restype = i32
        L.orc_lang_id.argtypes = [ctypes.c_char_p]
        L.orc_evaluate.restype = u64
        L.orc_evaluate.argtypes = [i32, _u8p, u64, _u64p]
        L.orc_trace_bff.restype = u64
        L.orc_trace_bff.argtypes = [i32, _u8p, u64, _u8p, u64, _u64p, _u64p]
        L.orc_init_programs.restype = None
        L.orc_init_programs.argtypes = [_u8p, u64, u64, i32]
        L.orc_pairing_shuffle.restype = None
        L.orc_pairing_shuffle.argtypes = [_u32p, u64, u64, u64]
        L.orc_pairing_fixed.restype = None
        L.orc_pairing_fixed.argtypes = [_u32p, u64, u64]
        L.orc_pairing_cyclic.restype = None
        L.orc_pairing_cyclic.argtypes = [_u32p, u64, u64]
        L.orc_pairing_2d.restype = u64
        L.orc_pairing_2d.argtypes = [_u32p, u64, _u64p, _u32p, u64, u64]
        L.orc_mutate_and_run.restype = u64
        L.orc_mutate_and_run.argtypes = [i32, _u8p, _u32p, u64, u64, u64, u32,
                                         _u32p, _u32p, i32]
        L.orc_pairing_repeats.restype = u64
        L.orc_pairing_repeats.argtypes = [_u32p, u64]
        L.orc_debug_repeats_seen.restype = u64
        L.orc_debug_repeats_seen.argtypes = []
        L.orc_check_selfrep.restype = None
        L.orc_check_selfrep.argtypes = [i32, _u8p, u64, u64, u64, _u64p, i32]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def lang_id(name: str) -> int:
    i = lib().orc_lang_id(name.encode())
    if i < 0:
        raise ValueError(f"unknown language {name}")
    return i


def splitmix64(x: int) -> int:
    return int(lib().orc_splitmix64(x & (2**64 - 1)))


def epoch_seed(params_seed: int, x: int) -> int:
    return int(lib().orc_epoch_seed(params_seed & (2**64 - 1), x & (2**64 - 1)))


def evaluate(lang: str, tape: np.ndarray, steps: int = 8192):
    """Runs in place on a uint8[128] copy; returns (final_tape, ops, steps)."""
    t = np.ascontiguousarray(tape, dtype=np.uint8).copy()
    assert t.size == 128
    st = ctypes.c_uint64(0)
    ops = lib().orc_evaluate(lang_id(lang), _p(t, _u8p), steps, ctypes.byref(st))
    return t, int(ops), int(st.value)


def trace_bff(lang: str, tape: np.ndarray, steps: int = 8192):
    """Step trace of a BFF-family tape: (states uint8[k, 140], ops, steps),
    each state = u32 pc, head0, head1 (little-endian) + the 128-byte tape
    before the step."""
    t = np.ascontiguousarray(tape, dtype=np.uint8).copy()
    assert t.size == 128
    out = np.zeros((max(steps, 1), 140), dtype=np.uint8)
    k, st = ctypes.c_uint64(0), ctypes.c_uint64(0)
    ops = lib().orc_trace_bff(lang_id(lang), _p(t, _u8p), steps, _p(out, _u8p), steps,
                              ctypes.byref(k), ctypes.byref(st))
    return out[:k.value].copy(), int(ops), int(st.value)


def init_programs(n: int, seed_value: int, zero_init: bool = False) -> np.ndarray:
    soup = np.zeros(n * 64, dtype=np.uint8)
    lib().orc_init_programs(_p(soup, _u8p), n, seed_value, int(zero_init))
    return soup


def pairing_shuffle(n: int, params_seed: int, epoch: int) -> np.ndarray:
    s = np.zeros(n, dtype=np.uint32)
    lib().orc_pairing_shuffle(_p(s, _u32p), n, params_seed, epoch)
    return s


def pairing_fixed(n: int, epoch: int) -> np.ndarray:
    s = np.zeros(n, dtype=np.uint32)
    lib().orc_pairing_fixed(_p(s, _u32p), n, epoch)
    return s


def pairing_cyclic(n: int, epoch: int) -> np.ndarray:
    s = np.zeros(n, dtype=np.uint32)
    lib().orc_pairing_cyclic(_p(s, _u32p), n, epoch)
    return s


def pairing_2d(n: int, row_off: np.ndarray, nbr: np.ndarray, params_seed: int,
               epoch: int):
    s = np.zeros(n, dtype=np.uint32)
    ro = np.ascontiguousarray(row_off, dtype=np.uint64)
    nb = np.ascontiguousarray(nbr, dtype=np.uint32)
    k = lib().orc_pairing_2d(_p(s, _u32p), n, _p(ro, _u64p), _p(nb, _u32p),
                             params_seed, epoch)
    return s, int(k)


def mutate_and_run(lang: str, soup: np.ndarray, s: np.ndarray, num_indices: int,
                   seed_value: int, mutation_prob: int, nthreads: int = 0):
    """In place on soup.  Returns (total_ops, pair_ops, pair_steps)."""
    n = soup.size // 64
    assert soup.dtype == np.uint8 and soup.flags.c_contiguous
    sa = np.ascontiguousarray(s, dtype=np.uint32)
    po = np.zeros(n // 2, dtype=np.uint32)
    ps = np.zeros(n // 2, dtype=np.uint32)
    tot = lib().orc_mutate_and_run(lang_id(lang), _p(soup, _u8p), _p(sa, _u32p),
                                   n, num_indices, seed_value, mutation_prob,
                                   _p(po, _u32p), _p(ps, _u32p), nthreads)
    return int(tot), po, ps


def pairing_repeats(s: np.ndarray, n: int) -> int:
    """Programs that sit in two or more different slots of a pairing."""
    sa = np.ascontiguousarray(s, dtype=np.uint32)
    return int(lib().orc_pairing_repeats(_p(sa, _u32p), n))


def debug_repeats_seen() -> int:
    """Repeats the debug build's per-epoch check has counted (0 otherwise)."""
    return int(lib().orc_debug_repeats_seen())


def check_selfrep(lang: str, soup: np.ndarray, seed_value: int,
                  count: int | None = None, nthreads: int = 0) -> np.ndarray:
    n = soup.size // 64
    count = n if count is None else count
    res = np.zeros(count, dtype=np.uint64)
    sp = np.ascontiguousarray(soup, dtype=np.uint8)
    lib().orc_check_selfrep(lang_id(lang), _p(sp, _u8p), n, count, seed_value,
                            _p(res, _u64p), nthreads)
    return res


class Soup:
    """Epoch loop of RunSimulation (common_language.h:344-597) on the oracle:
    init, per-epoch pairing, mutate+run.  Used by tests and the CPU baseline."""

    def __init__(self, lang: str, n: int, seed: int = 0,
                 mutation_prob: int = 1 << 18, zero_init: bool = False,
                 nthreads: int = 0):
        self.lang, self.n, self.seed, self.P = lang, n, seed, mutation_prob
        self.nthreads = nthreads
        self.soup = init_programs(n, epoch_seed(seed, 0), zero_init)
        self.epoch = 0

    def step(self, s: np.ndarray | None = None, num_indices: int | None = None):
        if s is None:
            s = pairing_shuffle(self.n, self.seed, self.epoch)
        if num_indices is None:
            num_indices = self.n // 2
        r = mutate_and_run(self.lang, self.soup, s, num_indices,
                           epoch_seed(self.seed, self.epoch), self.P,
                           self.nthreads)
        self.epoch += 1
        return r
