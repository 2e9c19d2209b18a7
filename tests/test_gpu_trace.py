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
    assert ops == ops_ref
    states = trace.read_states(path)
    assert len(states) == steps_ref
    assert states[0][3] == t.tobytes()
    with pytest.raises(ValueError):
        trace.evaluate_and_save(bytearray(64), path, 10)
    with pytest.raises(Exception):
        capi.trace_bff("forth", t.tobytes(), 10)
