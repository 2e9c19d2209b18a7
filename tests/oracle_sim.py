"""RunSimulation (common_language.h:344-597) replayed on the CPU oracle.

Test infrastructure: yields the callback-epoch records that the golden
fixtures hold (soup and pairing hashes, cumulative ops, self-rep scores), so
the oracle can be compared with the reference's recorded runs.
"""
from __future__ import annotations

import hashlib

import numpy as np

from oracle import oracle as O


def sha(a, n=None):
    h = hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return h[:n] if n else h


def bff_ascii_parse(text: str) -> np.ndarray:
    """Parse of the ASCII BFF map for command characters and '0' (the null
    byte), bff.inc.h:167-205; other characters are not needed here."""
    out = []
    for ch in text:
        if ch == "0":
            out.append(0)
        else:
            assert ch in "[]+-.,<>{}", ch
            out.append(ord(ch))
    return np.array(out, dtype=np.uint8)


def simulate(lang, n, seed, epochs, callback_interval=1, mutation_prob=1 << 18,
             fixed_shuffle=False, permute_programs=True, zero_init=False,
             reset_interval=None, initial_program=None, row_off=None, nbr=None,
             eval_selfrep=False, nthreads=0):
    """Runs epochs until a callback sees state.epoch >= epochs; returns one
    dict per callback, with the fields make_golden.py's run_epochs writes."""
    soup = O.init_programs(n, O.epoch_seed(seed, 0), zero_init)
    if initial_program is not None:
        parsed = bff_ascii_parse(initial_program)
        soup[:parsed.size] = parsed
    launched = n // 32 * 32
    reset_index = 1
    total_ops = 0
    rows = []
    epoch = 0
    while True:
        k = n // 2
        if row_off is not None:
            s, k = O.pairing_2d(n, row_off, nbr, seed, epoch)
        elif not permute_programs:
            s = O.pairing_cyclic(n, epoch)
        elif fixed_shuffle:
            s = O.pairing_fixed(n, epoch)
        else:
            s = O.pairing_shuffle(n, seed, epoch)
        ops, _, _ = O.mutate_and_run(lang, soup, s, k, O.epoch_seed(seed, epoch),
                                     mutation_prob, nthreads)
        total_ops += ops
        if epoch % callback_interval == 0:
            row = {"epoch": epoch + 1, "soup": sha(soup, 16), "shuffle": sha(s, 16),
                   "total_ops": total_ops}
            if eval_selfrep:
                rep = np.zeros(n, dtype=np.uint64)
                rep[:launched] = O.check_selfrep(lang, soup, O.epoch_seed(seed, epoch),
                                                 launched, nthreads)
                row["selfrep"] = sha(rep, 16)
                row["selfrep_ge5"] = int((rep >= 5).sum())
                row["selfrep_sum"] = int(rep.sum())
            rows.append(row)
            if epoch + 1 >= epochs:
                return rows, soup
        if reset_interval is not None and epoch % reset_interval == 0:
            # InitPrograms with num_programs = n over the first launched
            # programs (common_language.h:590-594).
            fresh = O.init_programs(n, O.epoch_seed(seed, reset_index), zero_init)
            soup[:launched * 64] = fresh[:launched * 64]
            reset_index += 1
        epoch += 1
