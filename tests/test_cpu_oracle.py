"""CPU checks of the oracle: known answers, restatements of the pairing rules,
and every golden fixture recorded from the compiled reference.

These pin the oracle that the GPU parity tests compare against.
"""
import json
import os
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.oracle_sim import sha, simulate
from tests.tapegen import sm_tapes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = (1 << 64) - 1


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module", autouse=True)
def _built(oracle):
    return oracle


def sm_py(x):
    """SplitMix64 output function (Vigna), common_language.h:131-136."""
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix64_known_answers():
    # First outputs of the published splitmix64 generator seeded with 0:
    # state 0 -> 0xe220a8397b1dcdaf, state 0x9e37.. -> 0x6e789e6aa1b965f4.
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert O.splitmix64(1) == 0x910A2DEC89025CC1
    for x in [2, 12345, 2**63, M64]:
        assert O.splitmix64(x) == sm_py(x)


def test_epoch_seed_composition():
    # seed(x) = SM(SM(params.seed) ^ SM(x)), common_language.h:367-369.
    for s, x in [(0, 0), (10248, 7), (2**64 - 1, 2**40 + 3)]:
        assert O.epoch_seed(s, x) == sm_py(sm_py(s) ^ sm_py(x))


def test_init_programs_formula():
    # InitPrograms (common_language.h:138-155): byte j of program i is
    # SM(64*N*seed + 64*i + j) & 255 (wrapping), so it depends on N.
    sv = O.epoch_seed(5, 0)
    for n in (32, 64):
        a = O.init_programs(n, sv)
        for i in (0, 1, 777, n * 64 - 1):
            assert int(a[i]) == sm_py((64 * n * sv + i) & M64) & 255
    assert not O.init_programs(8, sv, zero_init=True).any()


def fy_py(n, seed, epoch):
    s = list(range(n))
    for i in range(n - 1, -1, -1):
        j = sm_py(O.epoch_seed(seed, epoch * n + i)) % (i + 1)
        s[i], s[j] = s[j], s[i]
    return np.array(s, dtype=np.uint32)


@pytest.mark.parametrize("n", [2, 3, 10, 64, 257])
def test_shuffle_matches_sequential_fisher_yates(n):
    # common_language.h:420-425 restated with Python integers.
    for seed, e in [(0, 0), (10248, 5), (2**63, 99)]:
        s = O.pairing_shuffle(n, seed, e)
        assert np.array_equal(s, fy_py(n, seed, e))
        assert np.array_equal(np.sort(s), np.arange(n))


def test_fixed_and_cyclic_pairing_rules():
    n = 64
    for e in range(12):
        off = (1 << (e % 6 + 1)) - 1
        exp = np.array([((i * off) % n) ^ (e & 1) for i in range(n)], dtype=np.uint32)
        assert np.array_equal(O.pairing_fixed(n, e), exp)
        cyc = np.arange(n, dtype=np.uint32) if e % 2 else np.roll(np.arange(n, dtype=np.uint32), 1)
        assert np.array_equal(O.pairing_cyclic(n, e), cyc)


def test_pairing_2d_properties():
    n, seed = 200, 4
    rng = np.random.default_rng(0)
    rows = [sorted(set(rng.integers(0, n, size=rng.integers(0, 5)).tolist()) - {i})
            for i in range(n)]
    row_off = np.zeros(n + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([v for r in rows for v in r] or [0], dtype=np.uint32)
    for e in range(5):
        s, k = O.pairing_2d(n, row_off, nbr, seed, e)
        assert np.array_equal(np.sort(s), np.arange(n))
        for p in range(k):
            a, b = int(s[2 * p]), int(s[2 * p + 1])
            assert b in rows[a]
        # Unmatched programs follow in increasing order.
        assert np.all(np.diff(s[2 * k:].astype(np.int64)) > 0)


@pytest.mark.parametrize("lang", O.LANGS)
def test_evaluate_vs_reference_vectors(lang):
    g = load("eval_vectors.json")["langs"][lang]
    tapes = sm_tapes(lang, g["count"], g["seed"])
    ops, crc, finals = [], [], []
    for t in tapes:
        f, o, _ = O.evaluate(lang, t, g["steps"])
        ops.append(o)
        crc.append(zlib.crc32(f.tobytes()))
        finals.append(f)
    assert ops == g["ops"]
    assert crc == g["final_crc32"]
    assert sha(np.stack(finals)) == g["finals_sha256"]
    for steps, lim in g["limits_first64"].items():
        lo, lc = [], []
        for t in tapes[:64]:
            f, o, _ = O.evaluate(lang, t, int(steps))
            lo.append(o)
            lc.append(zlib.crc32(f.tobytes()))
        assert lo == lim["ops"], steps
        assert lc == lim["final_crc32"], steps


def _check_rows(rows, expected):
    assert len(rows) == len(expected)
    for r, x in zip(rows, expected):
        for key, val in x.items():
            assert r[key] == val, (key, r["epoch"])


@pytest.mark.parametrize("lang", O.LANGS)
def test_epochs_vs_reference_runs(lang):
    g = load("epochs_small.json")
    p = g["params"]
    rows, _ = simulate(lang, p["num_programs"], p["seed"], p["epochs"],
                       callback_interval=p["callback_interval"])
    _check_rows(rows, g["langs"][lang])


@pytest.mark.parametrize("lang", O.LANGS)
def test_selfrep_vs_reference_runs(lang):
    g = load("selfrep_small.json")
    p = g["params"]
    rows, _ = simulate(lang, p["num_programs"], p["seed"], p["epochs"],
                       callback_interval=p["callback_interval"], eval_selfrep=True)
    _check_rows(rows, g["langs"][lang])


def grid_pattern(w, h, r=2):
    """Neighbour lists in the order make_2d_pattern.py:17-29 writes them."""
    rows = []
    for i in range(h):
        for j in range(w):
            rows.append([y * w + x for y in range(i - r, i + r + 1)
                         for x in range(j - r, j + r + 1)
                         if 0 <= y < h and 0 <= x < w and (y, x) != (i, j)])
    row_off = np.zeros(len(rows) + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(x) for x in rows])
    return row_off, np.array([v for x in rows for v in x], dtype=np.uint32)


@pytest.mark.parametrize("case", ["bff_noheads_32x16", "forth_16x16"])
def test_pairing_2d_vs_reference_runs(case):
    c = load("pairing_2d.json")["cases"][case]
    row_off, nbr = grid_pattern(c["w"], c["h"], c["radius"])
    rows, _ = simulate(c["lang"], c["w"] * c["h"], c["seed"], len(c["rows"]),
                       row_off=row_off, nbr=nbr)
    _check_rows(rows, c["rows"])


@pytest.mark.parametrize("case", ["fixed_shuffle", "cyclic", "zero_init", "reset_interval_8",
                                  "initial_program", "mutation_1e-3"])
def test_modes_vs_reference_runs(case):
    c = load("modes.json")["cases"][case]
    rows, _ = simulate(c["lang"], c["num_programs"], c["seed"], len(c["rows"]), **c["kwargs"])
    _check_rows(rows, c["rows"])


def test_full_size_first_epoch_vs_reference():
    path = os.path.join(GOLDEN, "full_size.json")
    if not os.path.exists(path):
        pytest.skip("full_size.json not generated")
    g = load("full_size.json")
    p = g["params"]
    for seed, rows in g["seeds"].items():
        got, soup = simulate(p["lang"], p["num_programs"], int(seed), 1,
                             callback_interval=p["callback_interval"])
        assert got[0]["total_ops"] == rows[0]["total_ops"]
        assert sha(soup) == rows[0]["soup"]


def test_long_run_first_callbacks_vs_reference():
    # The oracle's first 65 epochs of the 1025-epoch bff_noheads run at 2^17
    # (seed 0) against the reference's recorded soup and pairing hashes.
    g = load("bff_noheads_2p17_seed0_1025_epochs.json")
    p = g["params"]
    rows, soup = simulate(p["lang"], p["num_programs"], p["seed"], 65,
                          callback_interval=p["callback_interval"])
    assert [r["epoch"] for r in rows] == [1, 65]
    for got, want in zip(rows, g["rows"][:2]):
        assert got["total_ops"] == want["total_ops"]
    assert sha(soup) == g["rows"][1]["soup"]


def test_oracle_vs_live_reference_build():
    """Where oracle/_ref is built (from the reference sources), evaluate and
    one mutate-and-run epoch agree on fresh random inputs."""
    from oracle import refimpl as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    from tests.tapegen import biased_tapes
    for lang in O.LANGS:
        for t in biased_tapes(lang, 200, 99):
            f, o, _ = O.evaluate(lang, t)
            rf, ro = R.evaluate(lang, t)
            assert o == ro and np.array_equal(f, rf), lang
        n = 512
        soup = O.init_programs(n, O.epoch_seed(3, 0))
        s = O.pairing_shuffle(n, 3, 0)
        ref_soup = soup.copy()
        O.mutate_and_run(lang, soup, s, n // 2, O.epoch_seed(3, 0), 1 << 18)
        R.mutate_and_run(lang, ref_soup, s, O.epoch_seed(3, 0), 1 << 18, n // 2)
        assert np.array_equal(soup, ref_soup), lang
