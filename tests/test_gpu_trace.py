"""GPU step traces (cubff_trace_bff) against the oracle's, byte for byte."""
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests.tapegen import biased_tapes

pytestmark = pytest.mark.gpu

BFF_LANGS = ["bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit", "bff_perm",
             "bff_selfmove"]


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


@pytest.mark.parametrize("lang", BFF_LANGS)
def test_trace_matches_oracle(capi, lang):
    tapes = biased_tapes(lang, 24, seed=11)
    for i, t in enumerate(tapes):
        for steps in (0, 1, 37, 8192):
            image, ops = capi.trace_bff(lang, t.tobytes(), steps)
            states, ops_ref, _ = O.trace_bff(lang, t, steps)
            header = b"BFF\0" + struct.pack("<II", 1, 128)
            assert image == header + states.tobytes(), (lang, i, steps)
            assert ops == ops_ref


def test_evaluate_and_save(capi, tmp_path):
    from cubff_amd import trace
    t = biased_tapes("bff", 1, seed=3)[0]
    path = str(tmp_path / "run.bff")
    ops = trace.evaluate_and_save(bytearray(t.tobytes()), path, 500)
    _, ops_ref, steps_ref = O.evaluate("bff", t, 500)
    _, _, steps_one_more = O.evaluate("bff", t, 501)
    assert ops == trace.reference_return(ops_ref, steps_ref, 500, steps_one_more)
    states = trace.read_states(path)
    assert len(states) == steps_ref
    assert states[0][3] == t.tobytes()
    # At the step cap the reference returns one less than Evaluate.
    loop = np.full(128, ord("a"), dtype=np.uint8)
    loop[0], loop[1], loop[2], loop[3] = 5, 5, ord("["), ord("]")
    _, ops_cap, steps_cap = O.evaluate("bff", loop, 300)
    assert steps_cap == 300
    assert trace.evaluate_and_save(bytearray(loop.tobytes()), path, 300) == ops_cap - 1
    assert len(trace.read_states(path)) == 300
    with pytest.raises(ValueError):
        trace.evaluate_and_save(bytearray(64), path, 10)
    with pytest.raises(Exception):
        capi.trace_bff("forth", t.tobytes(), 10)


LANGS = BFF_LANGS + ["forth", "forthcopy", "forthtrivial", "forthtrivial_reset", "subleq",
                     "rsubleq4"]


@pytest.mark.parametrize("lang", LANGS)
def test_step_records_match_oracle(capi, lang):
    # cubff_trace_steps: the final tape, ops and steps of the run equal the
    # oracle's Evaluate; record k holds the tape after k steps (the oracle
    # run with a step limit of k); BFF records equal the oracle's traces.
    tapes = biased_tapes(lang, 12, seed=5)
    for i, t in enumerate(tapes):
        for steps in (0, 1, 50, 8192):
            rec, final, ops, run = capi.trace_steps(lang, t.tobytes(), steps, cap=min(steps, 600))
            f_ref, ops_ref, st_ref = O.evaluate(lang, t, steps)
            assert np.array_equal(final, f_ref) and ops == ops_ref and run == st_ref, (lang, i)
            assert len(rec) == min(run, 600)
            for k in sorted({0, 1, 7, len(rec) // 2, len(rec) - 1}):
                if 0 <= k < len(rec):
                    assert np.array_equal(rec[k, 16:144], O.evaluate(lang, t, k)[0]), (lang, i, k)
            if lang in BFF_LANGS and len(rec):
                states, _, _ = O.trace_bff(lang, t, steps)
                n = len(rec)
                assert np.array_equal(rec[:, :12], states[:n, :12]), (lang, i)
                assert np.array_equal(rec[:, 16:144], states[:n, 12:]), (lang, i)
                assert not rec[:, 12:16].any() and not rec[:, 144:].any()
