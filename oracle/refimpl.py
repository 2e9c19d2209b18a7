"""ctypes view of the compiled reference (oracle/_ref/libcubff_ref.so).

TEST INFRASTRUCTURE ONLY.  oracle/_ref is built by oracle/ref.mk from the
reference sources; it exists in this build container (and travels with the
tree), and is used to pin the oracle and to write golden fixtures.  Tests that
need it skip when it has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libcubff_ref.so")
MAIN_PATH = os.path.join(HERE, "_ref", "cubff_ref_main")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class RefState(ctypes.Structure):
    _fields_ = [
        ("soup", _u8p),
        ("soup_size", ctypes.c_ulonglong),
        ("shuffle_idx", _u32p),
        ("replication_per_prog", _u64p),
        ("epoch", ctypes.c_ulonglong),
        ("total_ops", ctypes.c_ulonglong),
        ("ops_per_run", ctypes.c_float),
        ("brotli_size", ctypes.c_ulonglong),
        ("brotli_bpb", ctypes.c_float),
        ("h0", ctypes.c_float),
        ("higher_entropy", ctypes.c_float),
        ("mops_s", ctypes.c_float),
    ]


CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(RefState), ctypes.c_void_p)

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        L.ref_run_simulation.restype = ctypes.c_int
        L.ref_run_simulation.argtypes = [
            ctypes.c_char_p, ctypes.c_ulonglong, ctypes.c_ulonglong,
            ctypes.c_uint32, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_char_p,
            ctypes.c_char_p, ctypes.c_ulonglong, ctypes.c_char_p, _u64p, _u32p,
            CALLBACK, ctypes.c_void_p]
        L.ref_eval_parsed_selfrep_debug.restype = ctypes.c_ulonglong
        L.ref_eval_parsed_selfrep_debug.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_ulonglong,
                                                    ctypes.c_ulonglong, ctypes.c_ulonglong]
        L.ref_eval_parsed_selfrep.restype = ctypes.c_ulonglong
        L.ref_eval_parsed_selfrep.argtypes = [ctypes.c_char_p, _u8p,
                                              ctypes.c_ulonglong,
                                              ctypes.c_ulonglong,
                                              ctypes.c_ulonglong]
        L.ref_parse.restype = ctypes.c_ulonglong
        L.ref_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _u8p,
                                ctypes.c_ulonglong]
    return _lib


def _fn(prefix: str, lang: str):
    f = getattr(lib(), f"{prefix}{lang}")
    return f


def evaluate(lang: str, tape: np.ndarray, steps: int = 8192):
    t = np.ascontiguousarray(tape, dtype=np.uint8).copy()
    f = _fn("ref_eval_", lang)
    f.restype = ctypes.c_ulonglong
    f.argtypes = [_u8p, ctypes.c_ulonglong]
    ops = f(t.ctypes.data_as(_u8p), steps)
    return t, int(ops)


def check_selfrep(lang: str, soup: np.ndarray, seed_value: int, count=None):
    n = soup.size // 64
    count = n if count is None else count
    sp = np.ascontiguousarray(soup, dtype=np.uint8).copy()
    res = np.zeros(n, dtype=np.uint64)
    f = _fn("ref_selfrep_", lang)
    f.restype = None
    f.argtypes = [_u8p, ctypes.c_ulonglong, ctypes.c_ulonglong,
                  ctypes.c_ulonglong, _u64p]
    f(sp.ctypes.data_as(_u8p), seed_value, n, count, res.ctypes.data_as(_u64p))
    return res[:count]


def mutate_and_run(lang: str, soup: np.ndarray, s: np.ndarray, seed_value: int,
                   mutation_prob: int, num_indices: int):
    n = soup.size // 64
    sa = np.ascontiguousarray(s, dtype=np.uint32)
    f = _fn("ref_mutate_run_", lang)
    f.restype = ctypes.c_ulonglong
    f.argtypes = [_u8p, _u32p, ctypes.c_ulonglong, ctypes.c_uint32,
                  ctypes.c_ulonglong, ctypes.c_ulonglong]
    return int(f(soup.ctypes.data_as(_u8p), sa.ctypes.data_as(_u32p),
                 seed_value, mutation_prob, n, num_indices))


def run_simulation(lang: str, num_programs: int, seed: int, callback,
                   mutation_prob: int = 1 << 18, callback_interval: int = 1,
                   permute_programs: bool = True, fixed_shuffle: bool = False,
                   zero_init: bool = False, eval_selfrep: bool = False,
                   reset_interval: int = 0, load_from=None, save_to=None,
                   save_interval: int = 0, initial_program=None,
                   row_off=None, nbr=None):
    """Runs the reference RunSimulation; callback(state: RefState) -> bool."""
    def cb(p, _user):
        return 1 if callback(p.contents) else 0

    c = CALLBACK(cb)
    ro = nb = None
    if row_off is not None:
        ro_a = np.ascontiguousarray(row_off, dtype=np.uint64)
        nb_a = np.ascontiguousarray(nbr, dtype=np.uint32)
        ro, nb = ro_a.ctypes.data_as(_u64p), nb_a.ctypes.data_as(_u32p)
    enc = lambda x: None if x is None else x.encode()
    lib().ref_run_simulation(
        lang.encode(), num_programs, seed, mutation_prob, callback_interval,
        int(permute_programs), int(fixed_shuffle), int(zero_init),
        int(eval_selfrep), reset_interval, enc(load_from), enc(save_to),
        save_interval, enc(initial_program), ro, nb, c, None)


def soup_of(state: RefState) -> np.ndarray:
    return np.ctypeslib.as_array(state.soup, shape=(state.soup_size,)).copy()


def shuffle_of(state: RefState, n: int) -> np.ndarray:
    return np.ctypeslib.as_array(state.shuffle_idx, shape=(n,)).copy()


def replication_of(state: RefState, n: int) -> np.ndarray:
    return np.ctypeslib.as_array(state.replication_per_prog, shape=(n,)).copy()


def parse(lang: str, program: str) -> np.ndarray:
    """LanguageInterface::Parse of the reference build."""
    buf = np.zeros(4096, dtype=np.uint8)
    n = lib().ref_parse(lang.encode(), program.encode(), buf.ctypes.data_as(_u8p), buf.size)
    assert n <= buf.size
    return buf[:n].copy()


def eval_parsed_selfrep(lang: str, parsed: np.ndarray, epoch: int, seed: int,
                        debug: bool = False) -> int:
    """EvalParsedSelfrep (common_language.h:326-342) of the reference build;
    with debug, its per-trial printing goes to this process's stdout."""
    p = np.ascontiguousarray(parsed, dtype=np.uint8)
    f = lib().ref_eval_parsed_selfrep_debug if debug else lib().ref_eval_parsed_selfrep
    return int(f(lang.encode(), p.ctypes.data_as(_u8p), p.size, epoch, seed))
