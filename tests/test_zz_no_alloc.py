"""cubff_eval_tapes_noalloc: the allocation-free variant of cubff_eval_tapes
gives the same final tapes, ops and steps, call after call on the shared
device scratch (a second, smaller call must not see the first call's
tapes), and rejects counts beyond its fixed maximum."""
import numpy as np
import pytest

from tests.tapegen import biased_tapes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from cubff_amd import capi
    capi.load()
    return capi


@pytest.mark.parametrize("lang", ["bff_noheads", "forth", "subleq", "rsubleq4"])
def test_noalloc_matches_eval_tapes(capi, lang):
    big = biased_tapes(lang, capi.NOALLOC_MAX_TAPES, seed=41)
    small = biased_tapes(lang, 300, seed=42)
    for i, (tapes, steps) in enumerate(((big, 8192), (small, 777))):
        print("noalloc", lang, "call", i, len(tapes), flush=True)
        want = capi.eval_tapes(lang, tapes, steps)
        got = capi.eval_tapes_noalloc(lang, tapes, steps)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), (lang, i)
    # Empty input is a no-op; too many tapes is an error, not a fault.
    t, o, s = capi.eval_tapes_noalloc(lang, small[:0], 10)
    assert t.shape == (0, 128)
    too_many = np.zeros((capi.NOALLOC_MAX_TAPES + 1, 128), dtype=np.uint8)
    with pytest.raises(RuntimeError, match="no-alloc scratch"):
        capi.eval_tapes_noalloc(lang, too_many, 10)
