"""Seeded tape generators shared by the parity tests and golden fixtures.

Random tapes are mostly comments for the BFF family, so a fraction of the
bytes is drawn from each language's opcode alphabet to exercise every op,
bracket matching, head wrap, stack overflow and the SUBLEQ branches.

`sm_tapes` depends only on SplitMix64 (implemented here in numpy), so the
golden fixtures built from it do not depend on numpy's RNG streams.
"""
import numpy as np

BFF_ASCII = list(b"[]+-.,<>{}\0")
ALPHABETS = {
    "bff": BFF_ASCII, "bff_noheads": BFF_ASCII, "bff8": BFF_ASCII,
    "bff8_noheads": BFF_ASCII, "bff_noheads_4bit": list(range(16)),
    "bff_perm": list(range(11)), "bff_selfmove": list(range(7)),
    "subleq": list(range(0, 128, 3)) + [0x80, 0xff, 0x7f],
    "rsubleq4": [0, 1, 2, 3, 4, 0x7c, 0x7d, 0x7e, 0x7f, 0x80, 0xfc, 0xff],
}
FORTH_ALPHABET = list(range(14)) + [0x13, 0x40, 0x41, 0x7f, 0x80, 0x81, 0xc0,
                                    0xc1, 0xbf, 0xff]

_U = np.uint64


def splitmix64_np(x: np.ndarray) -> np.ndarray:
    """Vectorised SplitMix64 finaliser (wrapping uint64 arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(_U) + _U(0x9E3779B97F4A7C15)
        z = (z ^ (z >> _U(30))) * _U(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> _U(27))) * _U(0x94D049BB133111EB)
        return z ^ (z >> _U(31))


def alphabet(lang):
    return ALPHABETS.get(lang, FORTH_ALPHABET)


def sm_tapes(lang, count, seed):
    """count x 128 uint8 tapes from SplitMix64 streams: tape t has opcode
    density d_t = (SM(seed, t) mod 1024) / 1024; byte b is an opcode byte with
    probability d_t, else uniform."""
    a = np.array(alphabet(lang), dtype=np.uint64)
    t = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = _U(seed) * _U(0x100000000) + t * _U(1024)
        dens = splitmix64_np(base) % _U(1024)
        idx = base[:, None] + _U(1) + np.arange(128, dtype=np.uint64)[None, :]
    r = splitmix64_np(idx)
    take = (r >> _U(20)) % _U(1024) < dens[:, None]
    op = a[((r >> _U(40)) % _U(a.size)).astype(np.int64)]
    out = np.where(take, op, r & _U(0xFF)).astype(np.uint8)
    return out


def biased_tapes(lang, count, seed):
    """count x 128 uint8 tapes with per-tape opcode density (numpy RNG; used
    only for live comparisons, never for stored fixtures)."""
    rng = np.random.default_rng(seed)
    tapes = rng.integers(0, 256, size=(count, 128), dtype=np.uint8)
    dens = rng.random(count)
    mask = rng.random((count, 128)) < dens[:, None]
    a = np.array(alphabet(lang), dtype=np.uint8)
    tapes[mask] = a[rng.integers(0, a.size, size=int(mask.sum()))]
    return tapes
