"""GPU parity of the batched path: every soup of a batch equals a separate
oracle run with its own language, size, seed and mutation threshold, under
every launch shape, with and without graph replay."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


def _oracle_run(lang, n, seed, prob, epochs):
    ref = O.Soup(lang, n, seed, prob)
    hist = []
    for e in range(epochs):
        tot, po, ps = ref.step(O.pairing_shuffle(n, seed, e))
        hist.append((ref.soup.copy(), po.copy(), ps.copy(), tot))
    return hist


SPEC = [  # (lang, n, seed, prob)
    ("bff_noheads", 4096, 0, 10737),
    ("bff_noheads", 1024, 1, 262144),
    ("forth", 2048, 2, 1073742),
    ("bff", 512, 3, 262144),
    ("subleq", 256, 4, 262144),
    ("bff_noheads", 2048, 5, 0),
    ("rsubleq4", 1024, 6, 131072),
    ("forthtrivial_reset", 512, 7, 262144),
]


@pytest.mark.parametrize("wg,tpl,graph", [(256, 1, False), (64, 3, False), (128, 2, True),
                                          (256, 4, True)])
def test_batch_matches_oracle(capi, wg, tpl, graph):
    E = 6
    refs = [_oracle_run(l, n, s, p, E) for (l, n, s, p) in SPEC]
    with capi.Batch([x[0] for x in SPEC], [x[1] for x in SPEC]) as b:
        b.init([x[2] for x in SPEC], [x[3] for x in SPEC])
        b.set_launch(wg, tpl)
        cum = [0] * len(SPEC)
        for e in range(E):
            b.run_epochs(e, 1, capi.SCHED_PLAIN, use_graph=graph)
            for i, (l, n, s, p) in enumerate(SPEC):
                soup, po, ps, tot = refs[i][e]
                cum[i] += tot
                assert np.array_equal(b.read_shuffle(i), O.pairing_shuffle(n, s, e)), (i, e)
                assert np.array_equal(b.read_soup(i), soup), (i, e)
                ops, steps, work = b.read_pair_stats(i, with_work=True)
                assert np.array_equal(ops, po) and np.array_equal(steps, ps), (i, e)
                assert (work <= steps).all(), (i, e)  # executed commands
                assert b.read_counters(i)[:2] == (cum[i], (e + 1) * (n // 2)), (i, e)


def test_batch_graph_multi_epoch(capi):
    # A captured epoch replayed 10 times equals 10 direct epochs.
    spec = SPEC[:3]
    with capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as a, \
            capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as b:
        for x in (a, b):
            x.init([s[2] for s in spec], [s[3] for s in spec])
        a.run_epochs(0, 10, capi.SCHED_PLAIN, use_graph=True)
        b.run_epochs(0, 10, capi.SCHED_PLAIN, use_graph=False)
        for i in range(len(spec)):
            assert np.array_equal(a.read_soup(i), b.read_soup(i))
            assert a.read_counters(i) == b.read_counters(i)


def test_batch_autotune_keeps_results(capi):
    spec = SPEC[:2]
    with capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as a, \
            capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as b:
        for x in (a, b):
            x.init([s[2] for s in spec], [s[3] for s in spec])
        wg, tpl, table = a.autotune(0, 1)
        assert wg in (64, 128, 256) and 1 <= tpl <= 4 and all(t > 0 for t in table)
        # 12 tuning epochs, 24 where they run as two-epoch deferring segments.
        tuned = a.next_epoch()
        assert tuned in (12, 24)
        b.run_epochs(0, tuned)
        for i in range(len(spec)):
            assert np.array_equal(a.read_soup(i), b.read_soup(i))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_pairing_drawn_ahead(capi, graph):
    # Each epoch draws the next epoch's pairing beside it; runs that continue
    # where the last one stopped reuse it, runs that start elsewhere draw
    # their own first.  Soups and pairings must equal the oracle either way.
    spec = SPEC[:4]
    E = 9
    refs = [_oracle_run(l, n, s, p, E) for (l, n, s, p) in spec]
    with capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as b:
        b.init([x[2] for x in spec], [x[3] for x in spec])
        b.run_epochs(0, 4, capi.SCHED_PLAIN, use_graph=graph)
        b.run_epochs(4, 2, capi.SCHED_PLAIN, use_graph=graph)  # continues: epoch 6 drawn
        for i, (l, n, s, p) in enumerate(spec):
            assert np.array_equal(b.read_shuffle(i), O.pairing_shuffle(n, s, 5)), i
            assert np.array_equal(b.read_soup(i), refs[i][5][0]), i
        # Jump: soups of epoch 7 written back, the run starts at epoch 8.
        for i in range(len(spec)):
            b.write_soup(i, refs[i][7][0])
        b.run_epochs(8, 1, capi.SCHED_PLAIN, use_graph=graph)
        for i, (l, n, s, p) in enumerate(spec):
            assert np.array_equal(b.read_shuffle(i), O.pairing_shuffle(n, s, 8)), i
            assert np.array_equal(b.read_soup(i), refs[i][8][0]), i


def test_config3_full_size_matches_oracle(capi):
    # The 64-soup batch (seeds 0..63, sizes cycling 2^12, 2^14, 2^16, 2^17,
    # mutation log-spaced 1e-5..1e-3), graph-replayed: its first soup of each
    # size equals a separate oracle run after each of two epochs.
    S = 64
    ns = [[4096, 16384, 65536, 131072][k % 4] for k in range(S)]
    probs = [int(round(10 ** (-5 + 2 * k / 63) * (1 << 30))) for k in range(S)]
    check = (0, 1, 2, 3)
    refs = {k: O.Soup("bff_noheads", ns[k], k, probs[k], nthreads=16) for k in check}
    with capi.Batch(["bff_noheads"] * S, ns) as b:
        b.init(list(range(S)), probs)
        for e in range(2):
            print("config 3 epoch", e, flush=True)
            b.run_epochs(e, 1, capi.SCHED_PLAIN, use_graph=True)
            for k, ref in refs.items():
                tot, po, ps = ref.step()
                assert np.array_equal(b.read_shuffle(k), O.pairing_shuffle(ns[k], k, e)), (k, e)
                assert np.array_equal(b.read_soup(k), ref.soup), (k, e)
                ops, steps = b.read_pair_stats(k)
                assert np.array_equal(ops, po) and np.array_equal(steps, ps), (k, e)


@pytest.mark.parametrize("budget,sched", [(768, "plain"), (6, "plain"), (24, "learned")])
def test_batch_deferring_segments_match_oracle(capi, budget, sched):
    # Runs of several epochs go through deferring segments (the bff-family
    # soups: long interactions resolved beside later epochs; the others
    # through the interpreter launches).  A small command budget defers most
    # interactions and builds long chains of entries.  After every run the
    # soups, pairings, counters and the last epoch's per-slot counts equal
    # the oracle's.
    E = 7
    refs = [_oracle_run(l, n, s, p, E) for (l, n, s, p) in SPEC]
    mode = capi.SCHED_LEARNED if sched == "learned" else capi.SCHED_PLAIN
    with capi.Batch([x[0] for x in SPEC], [x[1] for x in SPEC]) as b:
        b.init([x[2] for x in SPEC], [x[3] for x in SPEC])
        b.set_defer_budget(budget)
        b.set_pipeline(capi.FLOW_DEFER)  # under both schedules
        print("deferring segments", budget, sched, flush=True)
        for first, count in ((0, 3), (3, 4)):
            b.run_epochs(first, count, mode)
            e = first + count - 1
            for i, (l, n, s, p) in enumerate(SPEC):
                soup, po, ps, _ = refs[i][e]
                assert np.array_equal(b.read_shuffle(i), O.pairing_shuffle(n, s, e)), (i, e)
                assert np.array_equal(b.read_soup(i), soup), (i, e)
                ops, steps = b.read_pair_stats(i)
                assert np.array_equal(ops, po) and np.array_equal(steps, ps), (i, e)
                tot = sum(refs[i][x][3] for x in range(e + 1))
                assert b.read_counters(i)[:2] == (tot, (e + 1) * (n // 2)), (i, e)
        deferred = b.defer_stats()
        print("deferred", deferred, flush=True)
        if budget < 100:  # these soups' interactions rarely run past 768 commands
            assert deferred > 0
