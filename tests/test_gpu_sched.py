"""GPU tests of the learned scheduler: the MLP's MFMA forward/backward against
the CPU reference (tests/mlp_ref.py), training reduces the loss and beats the
untrained model, and learned scheduling leaves every soup bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import mlp_ref as R
from tests.tapegen import biased_tapes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_mlp_forward_matches_cpu(capi):
    tapes = biased_tapes("bff_noheads", 300, seed=11)
    with capi.Mlp(512, init_seed=3) as m:
        params = m.get_params()
        y = m.forward(tapes)
    y_bf, _, _ = R.forward_backward(params, tapes, np.zeros(300), bf16_points=True)
    y_64, _, _ = R.forward_backward(params, tapes, np.zeros(300), bf16_points=False)
    # Same rounding points: fp32 accumulation order is the only difference.
    assert np.max(np.abs(y - y_bf)) < 2e-2 * max(1.0, np.max(np.abs(y_bf)))
    assert _rel(y, y_bf) < 5e-3
    # Against float64 without intermediate rounding: bf16 activations cost
    # about three significant bits.
    assert _rel(y, y_64) < 3e-2


def test_mlp_fused_predict_matches_cpu(capi):
    # 300 tapes: two full 128-tape blocks and a ragged one.
    tapes = biased_tapes("bff_noheads", 300, seed=21)
    tapes[:8] = 0                     # single-bin histograms (counts of 128)
    tapes[8:16] = np.arange(128)      # every byte once
    with capi.Mlp(512, init_seed=6) as m:
        params = m.get_params()
        y = m.predict(tapes)
        y_fwd = m.forward(tapes)
    ref = R.predict_folded(params, tapes)
    # Same rounding points; fp32 accumulation order is the only difference.
    assert _rel(y, ref) < 5e-3
    assert np.max(np.abs(y - ref)) < 2e-2 * max(1.0, np.max(np.abs(ref)))
    # The unfolded forward differs by the bf16 rounding of p = x E.
    assert _rel(y, y_fwd) < 3e-2


def test_mlp_predict_pieces_matches_assembled_tapes(capi):
    # The scheduler reads the next epoch's tapes as two 64-byte pieces of the
    # current stage: same predictions as on the tapes put together, for
    # pieces used once, twice or not at all (ragged count of 300 tapes).
    rng = np.random.default_rng(5)
    pieces = biased_tapes("bff_noheads", 320, seed=22).reshape(-1, 64)  # 640 pieces
    index = rng.integers(0, pieces.shape[0], size=600).astype(np.uint32)
    index[:40] = np.arange(40)  # a few in order
    tapes = np.concatenate([pieces[index[0::2]], pieces[index[1::2]]], axis=1)
    with capi.Mlp(512, init_seed=6) as m:
        params = m.get_params()
        y = m.predict_pieces(pieces, index)
        y_tapes = m.predict(tapes)
        with pytest.raises(RuntimeError, match="out of range"):
            bad = index.copy()
            bad[7] = pieces.shape[0]
            m.predict_pieces(pieces, bad)
    # The per-row sums of the eight wavefronts' partial products are added
    # in arrival order: equal up to fp32 reassociation.
    assert np.allclose(y, y_tapes, rtol=1e-5, atol=1e-5)
    assert _rel(y, R.predict_folded(params, tapes)) < 5e-3


def test_mlp_gradients_match_cpu(capi):
    tapes = biased_tapes("bff_noheads", 256, seed=12)
    targets = np.random.default_rng(0).uniform(0, 13, 256).astype(np.float32)
    with capi.Mlp(256, init_seed=4) as m:
        params = m.get_params()
        loss, g = m.gradients(tapes, targets)
    _, loss_bf, g_bf = R.forward_backward(params, tapes, targets, bf16_points=True)
    _, loss_64, g_64 = R.forward_backward(params, tapes, targets, bf16_points=False)
    assert abs(loss - loss_bf) < 1e-3 * max(1.0, loss_bf)
    for name, (a, b) in R.BLOCKS.items():
        assert _rel(g[a:b], g_bf[a:b]) < 2e-2, name
        assert _rel(g[a:b], g_64[a:b]) < 6e-2, name


def test_mlp_gradient_finite_difference(capi):
    # Directional derivative of the smooth (unrounded, float64) model along
    # the GPU gradient equals the GPU gradient's norm, to bf16 accuracy.
    tapes = biased_tapes("bff", 128, seed=13)
    targets = np.random.default_rng(1).uniform(0, 10, 128).astype(np.float32)
    with capi.Mlp(128, init_seed=5) as m:
        params = m.get_params().astype(np.float64)
        _, g = m.gradients(tapes, targets)
    d = g.astype(np.float64) / np.linalg.norm(g)
    h = 1e-4
    f = lambda p: R.forward_backward(p, tapes, targets, False, False)[1]
    fd = (f(params + h * d) - f(params - h * d)) / (2 * h)
    assert abs(fd - np.linalg.norm(g)) < 0.05 * np.linalg.norm(g)


def _soup_data(n=8192, epochs=24, seed=9):
    """(staged tape, ops) pairs of a bff_noheads soup, from the oracle."""
    from tests.tapegen import splitmix64_np
    S = O.Soup("bff_noheads", n, seed)
    X, Y = [], []
    for e in range(epochs):
        s = O.pairing_shuffle(n, seed, e)
        sp = S.soup.reshape(n, 64)
        t = np.concatenate([sp[s[0::2]], sp[s[1::2]]], axis=1).copy()
        sv = O.epoch_seed(seed, e)
        k = np.arange(n // 2, dtype=np.uint64)
        with np.errstate(over="ignore"):
            idx = ((np.uint64(n) * np.uint64(sv) + k)[:, None] * np.uint64(128)
                   + np.arange(128, dtype=np.uint64)[None, :])
        r = splitmix64_np(idx)
        mut = ((r >> np.uint64(8)) & np.uint64((1 << 30) - 1)) < np.uint64(1 << 18)
        t[mut] = (r[mut] & np.uint64(255)).astype(np.uint8)
        _, po, _ = S.step(s)
        X.append(t)
        Y.append(po)
    return np.concatenate(X), np.log2(np.concatenate(Y).astype(np.float32) + 1)


def test_training_reduces_loss_and_beats_untrained(capi):
    X, y = _soup_data()
    ntr = 65536
    with capi.Mlp(ntr, init_seed=6) as m:
        p0 = m.forward(X[ntr:ntr + 16384])
        losses = m.train(X[:ntr], y[:ntr], steps=60, lr=2e-3)
        p1 = m.forward(X[ntr:ntr + 16384])
    yt = y[ntr:ntr + 16384]
    assert losses[-1] < 0.7 * losses[0]
    mse0, mse1 = np.mean((p0 - yt) ** 2), np.mean((p1 - yt) ** 2)
    base = np.mean((yt - yt.mean()) ** 2)
    assert mse1 < mse0 and mse1 < base
    # The longest interactions concentrate in the top predicted decile.
    top = np.argsort(-p1)[: len(p1) // 10]
    long_ = yt >= np.log2(1001)
    assert long_[top].sum() >= 2 * 0.1 * long_.sum()


def test_learned_schedule_is_bit_exact(capi):
    spec = [("bff_noheads", 4096, 0, 262144), ("forth", 2048, 1, 262144),
            ("bff", 2048, 2, 1 << 20)]
    with capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as a, \
            capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as b:
        for x in (a, b):
            x.init([s[2] for s in spec], [s[3] for s in spec])
        a.run_epochs(0, 20, capi.SCHED_LEARNED, use_graph=False)
        b.run_epochs(0, 20, capi.SCHED_PLAIN)
        for i in range(len(spec)):
            assert np.array_equal(a.read_soup(i), b.read_soup(i))
            assert a.read_counters(i) == b.read_counters(i)
        trainings, loss, buckets = a.sched_stats()
        assert trainings == 2 and np.isfinite(loss)
        assert buckets.sum() == sum(n // 2 for _, n, _, _ in spec)
    with capi.Batch([x[0] for x in spec], [x[1] for x in spec]) as c:
        c.init([s[2] for s in spec], [s[3] for s in spec])
        c.run_epochs(0, 20, capi.SCHED_LEARNED, use_graph=True)
        for i in range(len(spec)):
            ref = O.Soup(spec[i][0], spec[i][1], spec[i][2], spec[i][3])
            for e in range(20):
                ref.step()
            assert np.array_equal(c.read_soup(i), ref.soup)


def test_learned_order_across_runs_writes_and_modes(capi):
    # The order of the next epoch is built during the current one and kept
    # across calls; a soup write, a plain run or a graph/no-graph switch in
    # between must not change results (the order only permutes evaluation,
    # and a stale one is rebuilt).  Same soups and counters as all-plain.
    spec = [("bff_noheads", 4096, 0, 262144), ("forth", 1024, 1, 262144),
            ("bff", 2048, 2, 1 << 20)]
    langs, ns = [x[0] for x in spec], [x[1] for x in spec]
    seeds, probs = [x[2] for x in spec], [x[3] for x in spec]
    X = biased_tapes("bff_noheads", 1024, seed=31).reshape(-1)[: 2048 * 64]
    with capi.Batch(langs, ns) as a, capi.Batch(langs, ns) as b:
        for x in (a, b):
            x.init(seeds, probs)
        a.run_epochs(0, 5, capi.SCHED_LEARNED, use_graph=True)
        a.run_epochs(5, 3, capi.SCHED_PLAIN)
        a.run_epochs(8, 4, capi.SCHED_LEARNED, use_graph=False)
        a.write_soup(2, X)
        a.run_epochs(12, 1, capi.SCHED_LEARNED)
        a.run_epochs(13, 6, capi.SCHED_LEARNED, use_graph=True)
        b.run_epochs(0, 12, capi.SCHED_PLAIN)
        b.write_soup(2, X)
        b.run_epochs(12, 7, capi.SCHED_PLAIN)
        for i in range(len(spec)):
            assert np.array_equal(a.read_soup(i), b.read_soup(i)), i
            assert a.read_counters(i) == b.read_counters(i), i
            assert np.array_equal(a.read_shuffle(i), b.read_shuffle(i)), i


def test_learned_buckets_cut_interpreter_time(capi):
    # The gain the trained network buys: on the 64-soup batch (seeds 0..63,
    # sizes cycling 2^12..2^17, mutation 1e-5..1e-3) the interpreter launches
    # of the learned schedule take less device time than the plain ones over
    # the same epochs, with identical soups, and so do whole epochs: the
    # next epoch's inference runs beside the interpreter.  Two tapes per
    # lane, the shape the autotuner picks for this batch (with one, the
    # interpreter's workgroups leave the inference no room and whole epochs
    # are slower).  Both arms run the interpreter rounds, the method the
    # order serves (plain runs default to deferring segments instead).
    S = 64
    ns = [[4096, 16384, 65536, 131072][k % 4] for k in range(S)]
    probs = [int(round(10 ** (-5 + 2 * k / 63) * (1 << 30))) for k in range(S)]
    interp = {}
    total = {}
    soups = {}
    for name, sched in (("plain", capi.SCHED_PLAIN), ("learned", capi.SCHED_LEARNED)):
        print("timing", name, flush=True)
        with capi.Batch(["bff_noheads"] * S, ns) as b:
            b.init(list(range(S)), probs)
            b.set_launch(64, 2)
            b.set_pipeline(capi.FLOW_OFF)
            b.run_epochs(0, 8, sched)
            tot, it = b.run_epochs_timed(8, 12, sched)
            interp[name] = it
            total[name] = tot
            soups[name] = [b.read_soup(i) for i in range(S)]
    assert len(soups["plain"]) == S
    for i, (x, y) in enumerate(zip(soups["plain"], soups["learned"])):
        assert np.array_equal(x, y), i
    assert interp["learned"] < 0.95 * interp["plain"], interp
    assert total["learned"] < total["plain"], total


def _config5():
    """32 soups of 2^16 programs, seeds 100..131, bff on even seeds and forth
    on odd ones, default mutation."""
    seeds = list(range(100, 132))
    return (["bff" if x % 2 == 0 else "forth" for x in seeds], [65536] * 32, seeds,
            [1 << 18] * 32)


def test_config5_learned_with_graph_is_bit_exact(capi):
    # The learned scheduler's own workload, across the retraining at epoch
    # 16: graph-replayed learned epochs give every one of the 32 soups the
    # soup, counters, pairing and last per-pair results of plain epochs, and
    # the first three epochs of one bff and one forth soup are the oracle's.
    langs, ns, seeds, probs = _config5()
    E = 22
    refs = {i: O.Soup(langs[i], ns[i], seeds[i], probs[i], nthreads=16) for i in (0, 1)}
    out = {}
    for name, sched in (("learned", capi.SCHED_LEARNED), ("plain", capi.SCHED_PLAIN)):
        print("config 5", name, flush=True)
        with capi.Batch(langs, ns) as b:
            b.init(seeds, probs)
            if sched == capi.SCHED_LEARNED:
                for e in range(3):
                    b.run_epochs(e, 1, sched, use_graph=True)
                    for i, ref in refs.items():
                        tot, po, ps = ref.step()
                        assert np.array_equal(b.read_soup(i), ref.soup), (langs[i], e)
                        ops, steps = b.read_pair_stats(i)
                        assert np.array_equal(ops, po) and np.array_equal(steps, ps), (i, e)
                b.run_epochs(3, E - 3, sched, use_graph=True)
                trainings, loss, buckets = b.sched_stats()
                assert trainings >= 2 and np.isfinite(loss)
                assert buckets.sum() == sum(ns) // 2
            else:
                b.run_epochs(0, E, sched, use_graph=False)
            out[name] = [(b.read_soup(i), b.read_counters(i), b.read_shuffle(i),
                          b.read_pair_stats(i)) for i in range(len(ns))]
    for i in range(len(ns)):
        (sa, ca, ha, pa), (sb, cb, hb, pb) = out["learned"][i], out["plain"][i]
        assert np.array_equal(sa, sb), i
        assert ca == cb, i
        assert np.array_equal(ha, hb) and np.array_equal(ha, O.pairing_shuffle(ns[i], seeds[i], E - 1))
        assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1]), i
