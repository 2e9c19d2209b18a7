"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Bit-exact on every byte of every tape/soup, every per-slot ops/steps value,
every pairing and every self-rep score.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.tapegen import biased_tapes

pytestmark = pytest.mark.gpu

LANGS = O.LANGS


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("lang", LANGS)
def test_eval_tapes_vs_oracle(capi, lang):
    tapes = biased_tapes(lang, 3000, seed=hash(lang) % 2**32)
    g_t, g_ops, g_steps = capi.eval_tapes(lang, tapes)
    for i in range(tapes.shape[0]):
        f, ops, steps = O.evaluate(lang, tapes[i])
        assert ops == g_ops[i] and steps == g_steps[i], (lang, i, tapes[i].tobytes().hex())
        assert np.array_equal(f, g_t[i]), (lang, i, tapes[i].tobytes().hex())


def test_eval_tapes_step_limits(capi):
    # A tight loop that never ends: every step limit must be honoured exactly.
    tape = np.zeros(128, dtype=np.uint8)
    tape[0:3] = list(b"+[]")
    for steps in [0, 1, 2, 3, 100, 8191, 8192, 32768]:
        g_t, g_ops, g_st = capi.eval_tapes("bff_noheads", tape[None, :], steps=steps)
        f, ops, st = O.evaluate("bff_noheads", tape, steps)
        assert (g_ops[0], g_st[0]) == (ops, st), steps
        assert np.array_equal(g_t[0], f)


@pytest.mark.parametrize("lang", ["bff8", "bff8_noheads", "bff_selfmove", "subleq", "rsubleq4"])
def test_static_cycle_skip_vs_oracle(capi, lang):
    # Runs that end in a cycle on a settled tape are skipped whole periods at
    # a time; the results at any step limit must equal the step-by-step
    # oracle's.  Tapes whose oracle run reaches 8192 steps, at limits that
    # cut periods at arbitrary points.
    from tests.tapegen import biased_tapes
    long = []
    for seed in range(17, 27):  # SUBLEQ's capped runs are rare: draw until enough
        long += [t for t in biased_tapes(lang, 1500, seed=seed)
                 if O.evaluate(lang, t, 8192)[2] >= 8192]
        if len(long) >= 12:
            break
    long = long[:40]
    assert len(long) >= 5, len(long)
    batch = np.stack(long)
    for steps in (8192, 8191, 5003, 4096, 1000, 37, 30000):
        g_t, g_ops, g_st = capi.eval_tapes(lang, batch, steps=steps)
        for i, t in enumerate(long):
            f, ops, st = O.evaluate(lang, t, steps)
            assert (g_ops[i], g_st[i]) == (ops, st), (lang, i, steps)
            assert np.array_equal(g_t[i], f), (lang, i, steps)


@pytest.mark.parametrize("lang", LANGS)
def test_epochs_shuffle_vs_oracle(capi, lang):
    n, seed, P = 2048, 10248, 1 << 18
    ref = O.Soup(lang, n, seed, P)
    with capi.Soup(lang, n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        assert np.array_equal(g.read_soup(), ref.soup)
        cum = 0
        for e in range(12):
            g.run_epochs(capi.PAIR_SHUFFLE, seed, e, 1, P)
            s_ref = O.pairing_shuffle(n, seed, e)
            tot, po, ps = ref.step(s_ref)
            cum += tot
            s, k = g.read_shuffle()
            assert k == n // 2
            assert np.array_equal(s, s_ref), (lang, e)
            ops, steps = g.read_pair_ops()
            assert np.array_equal(ops, po), (lang, e)
            assert np.array_equal(steps, ps), (lang, e)
            assert np.array_equal(g.read_soup(), ref.soup), (lang, e)
            assert g.read_counters() == (cum, (e + 1) * (n // 2))


@pytest.mark.parametrize("n", [2, 6, 64, 1000, 4096, 65536, 131072])
def test_shuffle_pairing_sizes(capi, n):
    with capi.Soup("bff_noheads", n) as g:
        for seed, e in [(0, 0), (0, 7), (10248, 3), (2**63 + 5, 123456789)]:
            g.pair(capi.PAIR_SHUFFLE, seed, e)
            s, k = g.read_shuffle()
            assert np.array_equal(s, O.pairing_shuffle(n, seed, e)), (n, seed, e)


@pytest.mark.parametrize("mode", ["fixed", "cyclic"])
def test_fixed_and_cyclic_pairing(capi, mode):
    n, seed, P = 1024, 3, 1 << 18
    ref = O.Soup("bff_noheads", n, seed, P)
    m = capi.PAIR_FIXED if mode == "fixed" else capi.PAIR_CYCLIC
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        for e in range(24):
            g.run_epochs(m, seed, e, 1, P)
            s_ref = O.pairing_fixed(n, e) if mode == "fixed" else O.pairing_cyclic(n, e)
            ref.step(s_ref)
            s, _ = g.read_shuffle()
            assert np.array_equal(s, s_ref), e
            assert np.array_equal(g.read_soup(), ref.soup), e


def grid_pattern(w, h, r=2):
    """Neighbour lists in the order make_2d_pattern.py:17-29 writes them."""
    rows = []
    for i in range(h):
        for j in range(w):
            lst = []
            for ii in range(-r, r + 1):
                for jj in range(-r, r + 1):
                    y, x = i + ii, j + jj
                    if 0 <= y < h and 0 <= x < w and (ii or jj):
                        lst.append(y * w + x)
            rows.append(lst)
    row_off = np.zeros(len(rows) + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(x) for x in rows])
    return row_off, np.array([v for x in rows for v in x], dtype=np.uint32)


@pytest.mark.parametrize("w,h", [(32, 16), (64, 64)])
def test_pairing_2d_vs_oracle(capi, w, h):
    n, seed, P = w * h, 1, 1 << 18
    row_off, nbr = grid_pattern(w, h)
    ref = O.Soup("bff_noheads", n, seed, P)
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        g.set_interactions(row_off, nbr)
        for e in range(8):
            g.run_epochs(capi.PAIR_2D, seed, e, 1, P)
            s_ref, k_ref = O.pairing_2d(n, row_off, nbr, seed, e)
            ref.step(s_ref, k_ref)
            s, k = g.read_shuffle()
            assert k == k_ref, e
            assert np.array_equal(s, s_ref), e
            assert np.array_equal(g.read_soup(), ref.soup), e


def test_pairing_2d_partial_pattern(capi):
    # Programs without neighbours are skipped and left over in order.
    n, seed = 256, 9
    rng = np.random.default_rng(5)
    rows = [list(rng.integers(0, n, size=rng.integers(0, 4))) if i % 3 else []
            for i in range(n)]
    row_off = np.zeros(n + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([v for r in rows for v in r], dtype=np.uint32)
    with capi.Soup("bff", n) as g:
        g.set_interactions(row_off, nbr)
        for e in range(5):
            g.pair(capi.PAIR_2D, seed, e)
            s, k = g.read_shuffle()
            s_ref, k_ref = O.pairing_2d(n, row_off, nbr, seed, e)
            assert k == k_ref and np.array_equal(s, s_ref), e


@pytest.mark.parametrize("lang", LANGS)
def test_selfrep_vs_oracle(capi, lang):
    n, seed = 512, 4
    ref = O.Soup(lang, n, seed)
    for e in range(6):
        ref.step()
    sv = O.epoch_seed(seed, 6)
    with capi.Soup(lang, n) as g:
        g.write_soup(ref.soup)
        got = g.selfrep(sv)
    want = O.check_selfrep(lang, ref.soup, sv)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("lang", ["bff_noheads", "bff_perm"])
def test_selfrep_two_chunks_vs_oracle(capi, lang):
    # More programs than one chunk of the levelled check (set to 2^18 here),
    # with copy loops planted around the chunk boundary and elsewhere so that
    # chains reach every level (long runs, cycles, programs that copy
    # themselves); then the same count in one chunk.
    n, count = 1 << 19, (1 << 18) + 3000
    soup = O.init_programs(n, O.epoch_seed(11, 0))
    plants = [b"[[{.>]-]                ]", b"[<,]", b"[{.>>}]", b"[[[-]+]", b"{[.>]}" * 6,
              b"[>,<]" * 4]
    rng = np.random.default_rng(5)
    spots = [(1 << 18) - 2, (1 << 18) - 1, 1 << 18, (1 << 18) + 1, 0, 77] + \
        [int(x) for x in rng.integers(0, count, 400)]
    for k, q in enumerate(spots):
        prog = plants[k % len(plants)]
        off = q * 64 + int(rng.integers(0, 64 - len(prog) + 1))
        soup[off:off + len(prog)] = list(prog)
    sv = O.epoch_seed(11, 9)
    print("selfrep two chunks", lang, flush=True)
    with capi.Soup(lang, n) as g:
        g.write_soup(soup)
        g.set_selfrep_chunk(1 << 18)
        got = g.selfrep(sv, count)
        again = g.selfrep(sv, 4096)
        g.set_selfrep_chunk(0)
        whole = g.selfrep(sv, count)
    want = O.check_selfrep(lang, soup, sv, count)
    assert np.array_equal(got, want)
    assert np.array_equal(again, want[:4096])
    assert np.array_equal(whole, want)


def test_eval_selfrep_single(capi):
    prog = bytes(b"[[{.>]-]                ]") + bytes(20)
    for lang in ["bff", "bff_noheads"]:
        for epoch, seed in [(0, 0), (5, 17)]:
            got = capi.eval_selfrep(lang, prog, epoch, seed)
            p = np.zeros(64, dtype=np.uint8)
            p[: len(prog)] = list(prog)
            want = O.check_selfrep(lang, p, O.epoch_seed(seed, epoch))[0]
            assert got == want


def test_full_size_golden_sha(capi):
    """bff_noheads, 2^17 programs, seed 10248: soup SHA-256 after epochs 1 and
    65 equal the values recorded from the reference build."""
    n, seed, P = 131072, 10248, 1 << 18
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 1, P)
        assert _sha(g.read_soup()) == (
            "bfd2afbafb764feac7c3465b407b384181b99743b8c3a5b93cdd29e31c49666c")
        g.run_epochs(capi.PAIR_SHUFFLE, seed, 1, 64, P)
        assert _sha(g.read_soup()) == (
            "72826732daa298a988a3ffc30a6734dbfd406084472df72ea713919365d80c89")


@pytest.mark.parametrize("n", [2, 4096, 131072])
def test_byte_histogram_matches_soup(capi, n):
    # The callback block's h0 and byte tables come from this histogram.
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(3, 0))
        g.run_epochs(capi.PAIR_SHUFFLE, 3, 0, 3, 1 << 20)
        soup = g.read_soup()
        assert np.array_equal(g.byte_histogram(), np.bincount(soup, minlength=256).astype(np.uint64))
