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

import numpy as np

from cubff_amd import capi

_MAGIC = b"BFF\0"
_TAPE = 128


def evaluate_and_save(tape: bytes | bytearray, file_path: str, stepcount: int,
                      debug: bool = False, lang: str = "bff") -> int:
    """Runs the tape for at most `stepcount` steps, writes every state before
    a step to `file_path`, and returns what the reference returns: the
    executed non-no-op steps, less one when the step cap ended the run (see
    reference_return).  The tape passed in is not modified."""
    if len(tape) != _TAPE:
        raise ValueError(f"Program must be exactly {_TAPE} bytes long")
    image, ops = capi.trace_bff(lang, bytes(tape), stepcount)
    with open(file_path, "wb") as f:
        f.write(image)
    if debug:
        for pc, h0, h1, t in _states(image):
            print(f"pc={pc} head0={h0} head1={h1} tape={t.hex()}")
    steps = (len(image) - 12) // (12 + _TAPE)
    one_more = steps
    if stepcount > 0 and steps == stepcount:
        t = np.frombuffer(bytes(tape), dtype=np.uint8)
        one_more = int(capi.eval_tapes(lang, t, stepcount + 1)[2][0])
    return reference_return(ops, steps, stepcount, one_more)


def reference_return(ops: int, steps: int, stepcount: int, steps_one_more: int) -> int:
    """evaluate_and_save's return value from Evaluate's results.

    ops and steps are Evaluate's result and loop iterations at `stepcount`
    (bff.inc.h:356-389), steps_one_more the iterations of the same tape at
    stepcount + 1.  The reference's loop, `for i in range(stepcount)` with
    `i += 1; break` where the program ends (bff_interpreter.py:373-391),
    leaves i = stepcount - 1 when all stepcount steps run and the program is
    still going, and returns i - nskip: one less than Evaluate, whose i
    reaches stepcount (-1 when every step was a no-op).  A run that ends
    before the cap, or at its last step, returns Evaluate's value."""
    if stepcount > 0 and steps == stepcount and steps_one_more > stepcount:
        return ops - 1
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
