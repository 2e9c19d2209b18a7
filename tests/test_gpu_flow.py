"""GPU parity of the multi-epoch dataflow kernel (cubff_run_epochs with
count >= 2): every interaction starts once the interactions of the previous
epoch that wrote its programs are done, and the result must equal the
epoch-by-epoch loop bit for bit -- soups, the last epoch's pairing and
per-pair results, and the interval counters -- for every language, pairing
mode, soup size and segment split.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# The dataflow kernel (FLOW_ON) or the deferring epoch kernels (FLOW_DEFER)
# are forced in every check below; the default (FLOW_AUTO) only picks them
# where they measured faster.


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _pairing(capi, mode, n, seed, e, graph=None):
    if mode == capi.PAIR_2D:
        return O.pairing_2d(n, graph[0], graph[1], seed, e)
    if mode == capi.PAIR_SHUFFLE:
        return O.pairing_shuffle(n, seed, e)
    if mode == capi.PAIR_FIXED:
        return O.pairing_fixed(n, e)
    return O.pairing_cyclic(n, e)


def _check_run(capi, lang, n, seed, P, mode, first, count, nthreads=0, pipeline=None,
               budget=None, graph=None):
    """count epochs from `first` in one call against the oracle's loop."""
    pipeline = capi.FLOW_ON if pipeline is None else pipeline
    print("flow:", lang, n, seed, mode, first, count, pipeline, budget, flush=True)
    ref = O.Soup(lang, n, seed, P, nthreads=nthreads)

    def pairing(e):
        p = _pairing(capi, mode, n, seed, e, graph)
        return p if mode == capi.PAIR_2D else (p, n // 2)

    with capi.Soup(lang, n) as g:
        g.set_pipeline(pipeline)
        if budget is not None:
            g.set_defer_budget(budget)
        if graph is not None:
            g.set_interactions(*graph)
        g.init_programs(O.epoch_seed(seed, 0))
        for e in range(first):  # oracle catches up; the GPU gets the soup
            ref.step(*pairing(e))
        g.write_soup(ref.soup)
        g.reset_counters()
        tot = pairs = 0
        for e in range(first, first + count):
            s_ref, k_ref = pairing(e)
            t, po, ps = ref.step(s_ref, k_ref)
            tot += t
            pairs += k_ref
        g.run_epochs(mode, seed, first, count, P)
        assert _sha(g.read_soup()) == _sha(ref.soup), (lang, n, mode, count)
        s, k = g.read_shuffle()
        assert k == k_ref and np.array_equal(s, s_ref)
        ops, steps = g.read_pair_ops()
        assert np.array_equal(ops, po) and np.array_equal(steps, ps)
        assert g.read_counters() == (tot, pairs)
        return g.flow_stats()[2]


@pytest.mark.parametrize("lang", O.LANGS)
def test_flow_every_language(capi, lang):
    _check_run(capi, lang, 2048, 10248, 1 << 18, capi.PAIR_SHUFFLE, 0, 12)


@pytest.mark.parametrize("mode", ["shuffle", "fixed", "cyclic"])
def test_flow_segments_and_modes(capi, mode):
    # 150 epochs = segments of 64, 64 and 22 epochs.
    m = {"shuffle": capi.PAIR_SHUFFLE, "fixed": capi.PAIR_FIXED, "cyclic": capi.PAIR_CYCLIC}[mode]
    _check_run(capi, "bff_noheads", 256, 7, 1 << 20, m, 5, 150)


@pytest.mark.parametrize("n", [2, 4, 6, 130])
def test_flow_tiny_soups(capi, n):
    # One or a few pairs per epoch: every item waits on the one before.
    _check_run(capi, "bff_noheads", n, 3, 1 << 22, capi.PAIR_SHUFFLE, 0, 70)
    _check_run(capi, "forth", n, 4, 1 << 22, capi.PAIR_SHUFFLE, 2, 9)


def test_flow_equals_epoch_loop_at_2p17(capi):
    n, seed, P = 131072, 1, 1 << 18
    print("flow vs loop at", n, flush=True)
    out = []
    for pipeline in (capi.FLOW_ON, capi.FLOW_OFF):
        with capi.Soup("bff_noheads", n) as g:
            g.set_pipeline(pipeline)
            g.init_programs(O.epoch_seed(seed, 0))
            g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 40, P)
            out.append((_sha(g.read_soup()), g.read_counters(), _sha(g.read_pair_ops()[0])))
    assert out[0] == out[1]
    _check_run(capi, "bff_noheads", n, seed, P, capi.PAIR_SHUFFLE, 0, 3, nthreads=16)


def test_flow_wait_limit_stops_cleanly(capi):
    # A zero wait limit: the first wave that finds no queued item raises the
    # abort flag, every wave leaves, and the host reports the error.  The
    # context stays usable.
    n, seed, P = 4096, 11, 1 << 18
    with capi.Soup("bff_noheads", n) as g:
        g.init_programs(O.epoch_seed(seed, 0))
        g.set_pipeline(capi.FLOW_ON, 0.0)
        with pytest.raises(RuntimeError, match="wait exceeded"):
            g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 8, P)
        g.set_pipeline(capi.FLOW_ON, 10.0)
    _check_run(capi, "bff_noheads", n, seed, P, capi.PAIR_SHUFFLE, 0, 8)


# Deferring epoch kernels + concurrent resolver.

@pytest.mark.parametrize("lang", O.LANGS)
def test_defer_every_language(capi, lang):
    # A budget of 16 commands defers many interactions and long chains.
    d = _check_run(capi, lang, 2048, 10248, 1 << 18, capi.PAIR_SHUFFLE, 0, 12,
                   pipeline=capi.FLOW_DEFER, budget=16)
    assert d > 0


@pytest.mark.parametrize("mode", ["shuffle", "fixed", "cyclic"])
def test_defer_segments_and_modes(capi, mode):
    m = {"shuffle": capi.PAIR_SHUFFLE, "fixed": capi.PAIR_FIXED, "cyclic": capi.PAIR_CYCLIC}[mode]
    _check_run(capi, "bff_noheads", 256, 7, 1 << 20, m, 5, 150, pipeline=capi.FLOW_DEFER,
               budget=8)


@pytest.mark.parametrize("n", [2, 4, 6, 130])
def test_defer_tiny_soups(capi, n):
    for budget in (1, 64):
        _check_run(capi, "bff_noheads", n, 3, 1 << 22, capi.PAIR_SHUFFLE, 0, 70,
                   pipeline=capi.FLOW_DEFER, budget=budget)
        _check_run(capi, "forth", n, 4, 1 << 22, capi.PAIR_SHUFFLE, 2, 9,
                   pipeline=capi.FLOW_DEFER, budget=budget)


def test_defer_equals_epoch_loop_at_2p17(capi):
    n, seed, P = 131072, 1, 1 << 18
    print("defer vs loop at", n, flush=True)
    out = []
    for pipeline in (capi.FLOW_DEFER, capi.FLOW_OFF):
        with capi.Soup("bff_noheads", n) as g:
            g.set_pipeline(pipeline)
            g.init_programs(O.epoch_seed(seed, 0))
            g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 40, P)
            out.append((_sha(g.read_soup()), g.read_counters(), _sha(g.read_pair_ops()[0]),
                        _sha(g.read_pair_ops()[1])))
            if pipeline == capi.FLOW_DEFER:
                assert g.flow_stats()[2] > 0
    assert out[0] == out[1]
    _check_run(capi, "bff_noheads", n, seed, P, capi.PAIR_SHUFFLE, 0, 3, nthreads=16,
               pipeline=capi.FLOW_DEFER)


def test_defer_resolver_serves_every_run(capi):
    # Runs of different lengths one after another (a short one, then one
    # with several chunks of Fisher-Yates pairings): each run's resolver
    # takes deferred interactions (none of them falls back to running every
    # interaction in the epoch kernels because its resolver left early), and
    # the soup equals the epoch-by-epoch loop's.
    n, seed, P = 131072, 2, 1 << 18
    print("defer runs of 6, 40 and 70 epochs at", n, flush=True)
    out = []
    for pipeline in (capi.FLOW_DEFER, capi.FLOW_OFF):
        with capi.Soup("bff_noheads", n) as g:
            g.set_pipeline(pipeline)
            g.init_programs(O.epoch_seed(seed, 0))
            first = 0
            for count in (6, 40, 70):
                d0 = g.flow_stats()[2]
                g.run_epochs(capi.PAIR_SHUFFLE, seed, first, count, P)
                first += count
                if pipeline == capi.FLOW_DEFER:
                    assert g.flow_stats()[2] > d0, count
            out.append((_sha(g.read_soup()), g.read_counters(), _sha(g.read_pair_ops()[0]),
                        _sha(g.read_shuffle()[0])))
    assert out[0] == out[1]


def test_defer_segment_boundary_at_2p18(capi):
    # 2^18 programs: deferring segments of 256 epochs (their entry lists
    # hold 16 bytes per program and epoch, up to 1 GiB), so 300 epochs cross
    # a segment boundary.  Same soup, counters and last results as the
    # epoch-by-epoch loop.
    n, seed, P = 1 << 18, 3, 1 << 18
    print("defer over a segment boundary at", n, flush=True)
    out = []
    for pipeline in (capi.FLOW_DEFER, capi.FLOW_OFF):
        with capi.Soup("bff_noheads", n) as g:
            g.set_pipeline(pipeline)
            g.init_programs(O.epoch_seed(seed, 0))
            g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 300, P)
            out.append((_sha(g.read_soup()), g.read_counters(), _sha(g.read_pair_ops()[0]),
                        _sha(g.read_pair_ops()[1])))
            if pipeline == capi.FLOW_DEFER:
                assert g.flow_stats()[2] > 0
    assert out[0] == out[1]


def test_defer_with_serialized_dispatch(tmp_path):
    # When kernels cannot run side by side (AMD_SERIALIZE_KERNEL=3, as under
    # some profilers), the resolver leaves after a short idle wait and the
    # epoch kernels run every interaction themselves: same results.
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "serial.py"
    script.write_text(
        "import sys, hashlib\n"
        f"sys.path.insert(0, {root!r})\n"
        "from cubff_amd import capi\n"
        "from oracle import oracle as O\n"
        "n, seed, P = 2048, 5, 1 << 20\n"
        "ref = O.Soup('bff_noheads', n, seed, P)\n"
        "for e in range(6):\n"
        "    ref.step()\n"
        "with capi.Soup('bff_noheads', n) as g:\n"
        "    g.set_pipeline(capi.FLOW_DEFER)\n"
        "    g.set_defer_budget(8)\n"
        "    g.init_programs(O.epoch_seed(seed, 0))\n"
        "    g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 6, P)\n"
        "    assert (g.read_soup() == ref.soup).all()\n"
        "print('serialized ok', flush=True)\n")
    env = dict(os.environ, AMD_SERIALIZE_KERNEL="3")
    p = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, timeout=110)
    print(p.stdout.decode()[-500:], p.stderr.decode()[-2000:])
    assert p.returncode == 0 and b"serialized ok" in p.stdout


def _partial_graph(n, seed):
    """Neighbour lists where every third program has none: the 2-D pairing
    then evaluates fewer than n/2 pairs and leaves the rest unpaired."""
    rng = np.random.default_rng(seed)
    rows = [list(rng.integers(0, n, size=rng.integers(1, 5))) if i % 3 else []
            for i in range(n)]
    row_off = np.zeros(n + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(r) for r in rows])
    return row_off, np.array([v for r in rows for v in r], dtype=np.uint32)


@pytest.mark.parametrize("graph", ["grid", "partial"])
@pytest.mark.parametrize("pipeline", ["loop", "defer"])
def test_2d_pairing_multi_epoch(capi, graph, pipeline):
    # 2-D pairings are drawn one epoch ahead like the closed-form ones; the
    # deferring kernels run slots past the evaluated pairs as mutation only.
    from cubff_amd.patterns import grid_pattern
    n = 4096 if graph == "grid" else 3000
    gr = grid_pattern(64, 64) if graph == "grid" else _partial_graph(n, 3)
    pl = capi.FLOW_OFF if pipeline == "loop" else capi.FLOW_DEFER
    for budget in ((None,) if pipeline == "loop" else (8, 256)):
        _check_run(capi, "bff_noheads", n, 5, 1 << 20, capi.PAIR_2D, 2, 11, pipeline=pl,
                   budget=budget, graph=gr)
    _check_run(capi, "forth", n, 6, 1 << 20, capi.PAIR_2D, 0, 5, pipeline=pl,
               budget=None if pipeline == "loop" else 16, graph=gr)


def test_fixed_pairing_of_non_power_of_two_soup_stays_on_the_epoch_loop(capi):
    # The fixed shuffle ((i * offset) % n) ^ flip repeats programs when n is
    # not a power of two (n = 6 at epoch 1: 1,2,1,2,1,2).  The reference's
    # library accepts it (its CLI refuses it, main.cc:233-237) and its slots
    # then race on the shared programs.  Here such runs take one launch per
    # epoch whatever the pipeline: the dataflow and deferring kernels, which
    # track one writer per program, must never see them.
    P = 1 << 22
    for n in (6, 1000):
        for lang in ("forth", "bff_noheads"):
            for pipeline in (capi.FLOW_ON, capi.FLOW_DEFER, capi.FLOW_AUTO):
                print("fixed non-power-of-two:", lang, n, pipeline, flush=True)
                with capi.Soup(lang, n) as g:
                    g.set_pipeline(pipeline)
                    g.init_programs(O.epoch_seed(9, 0))
                    g.run_epochs(capi.PAIR_FIXED, 9, 0, 8, P)
                    s, k = g.read_shuffle()
                    assert k == n // 2 and np.array_equal(s, O.pairing_fixed(n, 7))
                    items, _, deferred = g.flow_stats()
                    assert items == 0 and deferred == 0
    # Power-of-two soups still take the multi-epoch kernels.
    with capi.Soup("forth", 1024) as g:
        g.set_pipeline(capi.FLOW_ON)
        g.init_programs(O.epoch_seed(9, 0))
        g.run_epochs(capi.PAIR_FIXED, 9, 0, 8, 1 << 22)
        assert g.flow_stats()[0] > 0


def test_defer_resolver_leaving_at_once(capi):
    # With a zero idle time every resolver wavefront leaves at its first
    # look unless something is already deferred: it swaps the empty list
    # for the gone mark in one atomic step, and an epoch kernel whose
    # deferral lands on that mark runs the interaction itself.  Either way
    # the soup, counters and last results are the oracle's.
    n, seed, P = 131072, 2, 1 << 18
    for budget in (768, 16):
        print("resolver idle 0, budget", budget, flush=True)
        ref = O.Soup("bff_noheads", n, seed, P, nthreads=16)
        with capi.Soup("bff_noheads", n) as g:
            g.set_pipeline(capi.FLOW_DEFER)
            g.set_resolver_idle(0.0)
            g.set_defer_budget(budget)
            g.init_programs(O.epoch_seed(seed, 0))
            g.reset_counters()
            tot = 0
            for e in range(6):
                t, po, ps = ref.step()
                tot += t
            g.run_epochs(capi.PAIR_SHUFFLE, seed, 0, 6, P)
            assert _sha(g.read_soup()) == _sha(ref.soup)
            ops, steps = g.read_pair_ops()
            assert np.array_equal(ops, po) and np.array_equal(steps, ps)
            assert g.read_counters() == (tot, 6 * (n // 2))
