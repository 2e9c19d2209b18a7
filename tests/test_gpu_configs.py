"""GPU parity at the benchmark configurations' real sizes (a few epochs each,
against the CPU oracle): every interpreter at 2^17 programs, the 512 x 256
neighbour grid, and the 2^22-program soup with its self-rep scores."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("lang", ["bff", "bff8", "forth", "forthtrivial", "subleq", "rsubleq4"])
def test_interpreters_at_2p17(capi, lang):
    n, seed, P = 131072, 1, 1 << 18
    print("config: all interpreters,", lang, n, flush=True)
    ref = O.Soup(lang, n, seed, P)
    with capi.Soup(lang, n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        for e in range(2):
            g.run_epochs(capi.PAIR_SHUFFLE, seed, e, 1, P)
            ref.step()
            assert _sha(g.read_soup()) == _sha(ref.soup), (lang, e)


def test_grid_512x256(capi):
    w, h, seed, P = 512, 256, 1, 1 << 18
    n = w * h
    print("config: 2-D grid", w, h, flush=True)
    from cubff_amd.patterns import grid_pattern
    row_off, nbr = grid_pattern(w, h, 2)
    assert nbr.size == 3122724  # lines the generator writes for this grid
    ref = O.Soup("bff_noheads", n, seed, P)
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        g.set_interactions(row_off, nbr)
        for e in range(3):
            g.run_epochs(capi.PAIR_2D, seed, e, 1, P)
            s_ref, k_ref = O.pairing_2d(n, row_off, nbr, seed, e)
            ref.step(s_ref, k_ref)
            s, k = g.read_shuffle()
            assert k == k_ref and np.array_equal(s, s_ref), e
            assert _sha(g.read_soup()) == _sha(ref.soup), e


def test_large_soup_2p22(capi):
    n, seed = 1 << 22, 7
    P = round(0.00024 * (1 << 30))  # --mutation_prob 0.00024 (main.cc:226)
    assert P == 257698
    print("config: large soup", n, flush=True)
    ref = O.Soup("bff_noheads", n, seed, P)
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        for e in range(2):
            g.run_epochs(capi.PAIR_SHUFFLE, seed, e, 1, P)
            ref.step()
        assert _sha(g.read_soup()) == _sha(ref.soup)
        # Self-rep scores of the first 4096 programs (a program's score does
        # not depend on how many are scored).
        sv = O.epoch_seed(seed, 1)
        assert np.array_equal(g.selfrep(sv, 4096), O.check_selfrep("bff_noheads", ref.soup, sv,
                                                                    4096))
