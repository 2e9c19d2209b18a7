"""Step traces of BFF-family interactions for the reference's visualiser.

`evaluate_and_save` keeps the name, arguments, return value and file format
of python/bff_interpreter.py:evaluate_and_save (:327-393): the file holds
b"BFF\\0", u32 format version 1, u32 tape size 128, then, before every step,
u32 pc, u32 head0, u32 head1 (little-endian) and the 128 tape bytes.  The
steps run on the GPU (cubff_trace_bff); `read_states` reads such a file back
with the checks of read_and_display_states (:395-430).
"""
from __future__ import annotations

import struct

from cubff_amd import capi

_MAGIC = b"BFF\0"
_TAPE = 128


def evaluate_and_save(tape: bytes | bytearray, file_path: str, stepcount: int,
                      debug: bool = False, lang: str = "bff") -> int:
    """Runs the tape for at most `stepcount` steps, writes every state before
    a step to `file_path`, and returns the number of executed non-no-op
    steps.  The tape passed in is not modified."""
    if len(tape) != _TAPE:
        raise ValueError(f"Program must be exactly {_TAPE} bytes long")
    image, ops = capi.trace_bff(lang, bytes(tape), stepcount)
    with open(file_path, "wb") as f:
        f.write(image)
    if debug:
        for pc, h0, h1, t in _states(image):
            print(f"pc={pc} head0={h0} head1={h1} tape={t.hex()}")
    return ops


def _states(image: bytes):
    off = 12
    while off + 12 + _TAPE <= len(image):
        pc, h0, h1 = struct.unpack_from("<III", image, off)
        yield pc, h0, h1, image[off + 12: off + 12 + _TAPE]
        off += 12 + _TAPE


def read_states(file_path: str):
    """[(pc, head0, head1, tape bytes)] of a trace file."""
    with open(file_path, "rb") as f:
        image = f.read()
    if image[:4] != _MAGIC:
        raise ValueError(f"Invalid file format: Expected 'BFF\\0', got {image[:4]}")
    version = struct.unpack_from("<I", image, 4)[0]
    if version != 1:
        raise ValueError(f"Unsupported format version: {version}")
    size = struct.unpack_from("<I", image, 8)[0]
    if size != _TAPE:
        raise ValueError(f"Unexpected tape size: {size}, expected {_TAPE}")
    return list(_states(image))
