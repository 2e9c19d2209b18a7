"""CPU checks of the BFF step trace (the visualiser's file format).

The oracle's trace is pinned against orc_evaluate, itself pinned by the
reference build's vectors: the tape of state k equals the tape after k steps
of Evaluate, the number of states equals Evaluate's step count, and each
step changes at most the bytes under the heads recorded before it.
"""
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests.tapegen import biased_tapes

BFF_LANGS = ["bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit", "bff_perm",
             "bff_selfmove"]


@pytest.mark.parametrize("lang", BFF_LANGS)
def test_trace_prefixes_match_evaluate(lang):
    tapes = biased_tapes(lang, 40, seed=7)
    for t in tapes:
        states, ops, steps = O.trace_bff(lang, t, 300)
        _, ops_ref, steps_ref = O.evaluate(lang, t, 300)
        assert (ops, steps) == (ops_ref, steps_ref)
        assert len(states) == steps
        for k in range(0, len(states), 7):
            after_k, _, _ = O.evaluate(lang, t, k)
            assert np.array_equal(states[k, 12:], after_k), (lang, k)
        for k in range(len(states) - 1):
            pc, h0, h1 = struct.unpack("<III", states[k, :12].tobytes())
            assert pc < 128 and h0 < 128 and h1 < 128
            changed = np.nonzero(states[k, 12:] != states[k + 1, 12:])[0]
            assert set(changed.tolist()) <= {h0, h1}


def test_trace_initial_state():
    t = np.zeros(128, dtype=np.uint8)
    t[0], t[1] = 200, 7  # heads start at tape[0] % 128, tape[1] % 128, pc 2
    states, _, _ = O.trace_bff("bff", t, 3)
    assert struct.unpack("<III", states[0, :12].tobytes()) == (2, 72, 7)
    states, _, _ = O.trace_bff("bff_noheads", t, 3)
    assert struct.unpack("<III", states[0, :12].tobytes()) == (0, 0, 0)
    assert O.trace_bff("bff", t, 0)[0].shape == (0, 140)


def test_read_states_checks_header(tmp_path):
    from cubff_amd import trace
    good = b"BFF\0" + struct.pack("<II", 1, 128) + struct.pack("<III", 2, 3, 4) + bytes(128)
    (tmp_path / "a.bff").write_bytes(good)
    assert trace.read_states(str(tmp_path / "a.bff")) == [(2, 3, 4, bytes(128))]
    for bad in (b"BFX\0" + good[4:], good[:4] + struct.pack("<I", 2) + good[8:],
                good[:8] + struct.pack("<I", 64) + good[12:]):
        (tmp_path / "b.bff").write_bytes(bad)
        with pytest.raises(ValueError):
            trace.read_states(str(tmp_path / "b.bff"))


def _python_return(lang, t, cap):
    from cubff_amd.trace import reference_return
    _, ops, steps = O.evaluate(lang, t, cap)
    _, _, steps1 = O.evaluate(lang, t, cap + 1)
    return reference_return(ops, steps, cap, steps1)


def test_evaluate_and_save_return_at_the_step_cap():
    # Cases derived by hand from the reference's loop
    # (python/bff_interpreter.py:373-393: `for i in range(stepcount)`,
    # `i += 1; break` where the program ends, return i - nskip).
    # Comments only (no command ever executes): each step is a no-op; at a
    # cap of 10 the loop leaves i = 9 with nskip = 10, so -1; with the whole
    # tape run (cap 300) the last step ends it: i = 128, nskip = 128, so 0.
    comments = np.full(128, ord("a"), dtype=np.uint8)
    assert _python_return("bff_noheads", comments, 10) == -1
    assert _python_return("bff_noheads", comments, 300) == 0
    # 128 '+' commands, no heads: the program ends after its 128th step.
    plus = np.full(128, ord("+"), dtype=np.uint8)
    assert _python_return("bff_noheads", plus, 128) == 128  # ends at the last step
    assert _python_return("bff_noheads", plus, 127) == 126  # cut by the cap: i = 126
    assert _python_return("bff_noheads", plus, 500) == 128
    assert _python_return("bff_noheads", plus, 0) == 0      # the loop never runs
    # bff with heads: bytes 0 and 1 put both heads on byte 0 (behind the
    # code) and pc starts at 2, so 126 commands end the tape.
    heads = plus.copy()
    heads[0] = heads[1] = 0
    assert _python_return("bff", heads, 126) == 126
    assert _python_return("bff", heads, 125) == 124
    # An endless loop "[+]" with a nonzero byte under head 0: always capped.
    loop = np.full(128, ord("a"), dtype=np.uint8)
    loop[0], loop[1], loop[2] = ord("+"), ord("["), ord("]")
    for cap in (5, 64, 8192):
        _, ops, steps = O.evaluate("bff_noheads", loop, cap)
        assert steps == cap
        assert _python_return("bff_noheads", loop, cap) == ops - 1
