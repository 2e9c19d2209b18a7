"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in the build container after `make -f oracle/ref.mk` (it needs
oracle/_ref/, which is built from the reference sources):

    python tests/golden/make_golden.py [--only NAME ...]

Every fixture is data: inputs (or the seeds that generate them with
tests/tapegen.py / SplitMix64) and the reference's outputs.  Hashes are
SHA-256 of the raw little-endian bytes (first 16 hex digits where noted).
"""
from __future__ import annotations

import argparse
import base64
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refimpl as R  # noqa: E402
from tests.tapegen import sm_tapes  # noqa: E402

LANGS = [
    "bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit",
    "bff_perm", "bff_selfmove", "forth", "forthcopy", "forthtrivial",
    "forthtrivial_reset", "subleq", "rsubleq4",
]


def sha(a, n=None):
    h = hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return h[:n] if n else h


def dump(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
    print("wrote", path, os.path.getsize(path), "bytes")


def grid_pattern(w, h, r=2):
    """Neighbour lists in the order make_2d_pattern.py:17-29 writes them."""
    rows = []
    for i in range(h):
        for j in range(w):
            lst = []
            for ii in range(-r, r + 1):
                for jj in range(-r, r + 1):
                    y, x = i + ii, j + jj
                    if 0 <= y < h and 0 <= x < w and (ii or jj):
                        lst.append(y * w + x)
            rows.append(lst)
    row_off = np.zeros(len(rows) + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum([len(x) for x in rows])
    return row_off, np.array([v for x in rows for v in x], dtype=np.uint32)


def gen_eval(count=512):
    out = {}
    for li, lang in enumerate(LANGS):
        seed = 1000 + li
        tapes = sm_tapes(lang, count, seed)
        ops, crc = [], []
        finals = []
        for t in tapes:
            f, o = R.evaluate(lang, t)
            ops.append(o)
            crc.append(zlib.crc32(f.tobytes()))
            finals.append(f)
        limits = {}
        for steps in [0, 1, 2, 7, 100, 1000, 32768]:
            lo, lc = [], []
            for t in tapes[:64]:
                f, o = R.evaluate(lang, t, steps)
                lo.append(o)
                lc.append(zlib.crc32(f.tobytes()))
            limits[str(steps)] = {"ops": lo, "final_crc32": lc}
        out[lang] = {"seed": seed, "count": count, "steps": 8192, "ops": ops,
                     "final_crc32": crc, "finals_sha256": sha(np.stack(finals)),
                     "limits_first64": limits}
    dump("eval_vectors.json", {
        "generator": "tests/tapegen.py:sm_tapes(lang, count, seed)",
        "source": "reference Language::Evaluate via oracle/_ref", "langs": out})


def run_epochs(lang, n, seed, epochs, **kw):
    rows = []

    def cb(st):
        soup = R.soup_of(st)
        shuf = R.shuffle_of(st, n)
        row = {"epoch": st.epoch, "soup": sha(soup, 16), "shuffle": sha(shuf, 16),
               "total_ops": st.total_ops}
        if kw.get("eval_selfrep"):
            rep = R.replication_of(st, n)
            row["selfrep"] = sha(rep.astype(np.uint64), 16)
            row["selfrep_ge5"] = int((rep >= 5).sum())
            row["selfrep_sum"] = int(rep.sum())
        rows.append(row)
        return st.epoch >= epochs

    R.run_simulation(lang, n, seed, cb, **kw)
    return rows


def gen_epochs():
    out = {}
    for lang in LANGS:
        out[lang] = run_epochs(lang, 1024, 10248, 48, callback_interval=1)
    dump("epochs_small.json", {
        "params": {"num_programs": 1024, "seed": 10248, "mutation_prob": 262144,
                   "callback_interval": 1, "epochs": 48},
        "source": "reference RunSimulation via oracle/_ref", "langs": out})


def gen_selfrep():
    out = {}
    for lang in LANGS:
        out[lang] = run_epochs(lang, 512, 77, 33, callback_interval=8,
                               eval_selfrep=True)
    dump("selfrep_small.json", {
        "params": {"num_programs": 512, "seed": 77, "callback_interval": 8,
                   "eval_selfrep": True, "epochs": 33},
        "source": "reference RunSimulation via oracle/_ref", "langs": out})


def gen_2d():
    cases = {}
    for (w, h, lang, seed) in [(32, 16, "bff_noheads", 1), (16, 16, "forth", 2)]:
        row_off, nbr = grid_pattern(w, h)
        n = w * h
        rows = run_epochs(lang, n, seed, 12, callback_interval=1,
                          row_off=row_off, nbr=nbr)
        cases[f"{lang}_{w}x{h}"] = {"w": w, "h": h, "lang": lang, "seed": seed,
                                   "radius": 2, "rows": rows}
    dump("pairing_2d.json", {"source": "reference RunSimulation via oracle/_ref",
                             "cases": cases})


def gen_modes():
    cases = {}
    spec = {
        "fixed_shuffle": dict(fixed_shuffle=True),
        "cyclic": dict(permute_programs=False),
        "zero_init": dict(zero_init=True),
        "reset_interval_8": dict(reset_interval=8),
        "initial_program": dict(initial_program="[[{.>]-]0000]-]>.{[[<<+-0"),
        "mutation_1e-3": dict(mutation_prob=round(1e-3 * (1 << 30))),
    }
    for name, kw in spec.items():
        lang = "bff_noheads" if name != "initial_program" else "bff"
        rows = run_epochs(lang, 512, 21, 20, callback_interval=1, **kw)
        cases[name] = {"lang": lang, "num_programs": 512, "seed": 21,
                       "kwargs": kw, "rows": rows}
    dump("modes.json", {"source": "reference RunSimulation via oracle/_ref",
                        "cases": cases})


def gen_full_size():
    out = {}
    for seed in [10248, 0]:
        rows = run_epochs("bff_noheads", 131072, seed, 257, callback_interval=64)
        out[str(seed)] = [{"epoch": r["epoch"], "soup": None, "total_ops": r["total_ops"]}
                          for r in rows]
        # full hashes for these
        full = []

        def cb(st):
            full.append(sha(R.soup_of(st)))
            return st.epoch >= 257

        R.run_simulation("bff_noheads", 131072, seed, cb, callback_interval=64)
        for r, h in zip(out[str(seed)], full):
            r["soup"] = h
    dump("full_size.json", {
        "params": {"lang": "bff_noheads", "num_programs": 131072,
                   "mutation_prob": 262144, "callback_interval": 64},
        "source": "reference RunSimulation via oracle/_ref", "seeds": out})


def gen_long_run():
    """bff_noheads, 2^17 programs, seed 0, default mutation, callbacks every
    64 epochs up to epoch 1025 (the CLI's 1000-epoch run at print_interval
    64): full soup and pairing hashes and cumulative ops at each callback."""
    rows = []

    def cb(st):
        rows.append({"epoch": st.epoch, "soup": sha(R.soup_of(st)),
                     "shuffle": sha(R.shuffle_of(st, 131072)), "total_ops": st.total_ops,
                     "mops_s": None})
        return st.epoch >= 1025

    R.run_simulation("bff_noheads", 131072, 0, cb, callback_interval=64)
    for r in rows:
        del r["mops_s"]
    dump("bff_noheads_2p17_seed0_1025_epochs.json", {
        "params": {"lang": "bff_noheads", "num_programs": 131072, "seed": 0,
                   "mutation_prob": 262144, "callback_interval": 64, "epochs": 1025},
        "source": "reference RunSimulation via oracle/_ref", "rows": rows})


def gen_cli_logs():
    """CSV logs of the reference CLI (test.sh:30 settings) with this image's
    libbrotli, plus the reference's own testdata/ files for comparison."""
    out = {}
    tmp = tempfile.mkdtemp()
    for lang in LANGS:
        log = os.path.join(tmp, lang + ".txt")
        subprocess.run([R.MAIN_PATH, "--lang", lang, "--max_epochs", "256",
                        "--disable_output", "--log", log, "--seed", "10248"],
                       check=True)
        out[lang] = open(log).read()
    shutil.rmtree(tmp)
    ref_dir = os.path.join("/root/reference", "testdata")
    dst = os.path.join(HERE, "reference_testdata")
    os.makedirs(dst, exist_ok=True)
    for lang in LANGS:
        shutil.copyfile(os.path.join(ref_dir, lang + ".txt"),
                        os.path.join(dst, lang + ".txt"))
    dump("cli_logs_brotli109.json", {
        "command": "bin/main --lang L --max_epochs 256 --disable_output --log F --seed 10248",
        "brotli": "libbrotlienc 1.0.9 (this image)", "logs": out})


def run_programs():
    """Program texts for --run cases: a fixed BFF program, random command
    strings for the BFF maps, random hex for Forth and SUBLEQ."""
    rng = np.random.default_rng(2024)
    out = {}
    for lang in LANGS:
        if lang.startswith("bff"):
            texts = ["[[{.>]-]0000]-]>.{[[<<+-0"]
            for k in range(3):
                texts.append("".join(rng.choice(list("[]+-.,<>{}0"), size=20 + 30 * k)))
        else:
            texts = ["".join(f"{b:02x}" for b in rng.integers(0, 256, size=m))
                     for m in (8, 40, 100)]
        out[lang] = texts
    return out


def gen_cli_cases():
    """Reference CLI outputs that need no timing: --run transcripts, --help,
    argument errors, and deterministic logs / checkpoints / frames of small
    runs with every mode flag."""
    main = R.MAIN_PATH
    tmp = tempfile.mkdtemp()

    def run(args, cwd=tmp):
        p = subprocess.run([main] + args, cwd=cwd, capture_output=True)
        return p.returncode, p.stdout, p.stderr

    runs = {}
    for lang, texts in run_programs().items():
        rows = []
        for t in texts:
            for steps in (0, 1000):
                rc, out, _ = run(["--lang", lang, "--run", t, "--run_steps", str(steps)])
                assert rc == 0
                parsed = R.parse(lang, t)
                rows.append({"text": t, "steps": steps, "parsed": parsed.tobytes().hex(),
                             "stdout_zlib_b64": base64.b64encode(zlib.compress(out, 9)).decode()})
        runs[lang] = rows

    selfrep = {}
    for lang, texts in run_programs().items():
        # Programs longer than one 64-byte tape overflow the reference's
        # buffer (common_language.h:330-331), so only short ones are used.
        selfrep[lang] = [{"text": t, "epoch": e, "seed": s,
                          "result": R.eval_parsed_selfrep(lang, R.parse(lang, t), e, s)}
                         for t in texts if R.parse(lang, t).size <= 64
                         for (e, s) in ((0, 0), (5, 7))]

    # Self-replicators harvested from 257-epoch soups (131072 programs, seed
    # 10248) of the languages where they appear that early.
    for lang in ("forthtrivial", "bff_noheads_4bit", "forthtrivial_reset"):
        best = []

        def cb(st, lang=lang, best=best):
            if st.epoch < 257:
                return False
            rep = R.replication_of(st, 131072)
            soup = R.soup_of(st).reshape(-1, 64)
            seen = set()
            for i in np.argsort(-rep.astype(np.int64), kind="stable"):
                prog = soup[i].tobytes()
                if prog not in seen:
                    seen.add(prog)
                    best.append((int(rep[i]), soup[i].copy()))
                if len(best) == 3:
                    break
            return True

        R.run_simulation(lang, 131072, 10248, cb, callback_interval=256, eval_selfrep=True)
        for score, prog in best:
            for e, s in ((0, 0), (256, 10248), (5, 7)):
                selfrep[lang].append({"parsed": prog.tobytes().hex(), "epoch": e, "seed": s,
                                      "soup_score": score,
                                      "result": R.eval_parsed_selfrep(lang, prog, e, s)})

    rc, _, help_err = run(["--help"])
    errors = []
    for args in (["--nope"], ["--num"], ["--fixed_shuffle", "--num", "1000", "--lang", "bff"],
                 ["--save_interval", "100", "--lang", "bff"],
                 ["--clear_interval", "100", "--lang", "bff"],
                 ["--stopping_selfrep_count", "3", "--lang", "bff"],
                 ["--draw_to_2d", ".", "--num", "1000", "--lang", "bff"],
                 ["--draw_to_2d", ".", "--grid_width_2d", "7", "--num", "1000", "--lang", "bff"],
                 ["--interaction_pattern", "missing.txt", "--lang", "bff"],
                 ["--lang", "nope"]):
        rc, out, err = run(args)
        errors.append({"args": args, "rc": rc, "stdout": out.decode(), "stderr": err.decode()})

    pattern = "\n".join(f"{a} {b}" for a, nb in enumerate(_grid_lists(32, 32)) for b in nb)
    with open(os.path.join(tmp, "pattern.txt"), "w") as f:
        f.write(pattern + "\n")
    base = ["--disable_output", "--log", "log.csv", "--num", "1024", "--max_epochs", "96",
            "--print_interval", "16", "--save_interval", "32"]
    specs = {
        "shuffle_bff_noheads": ["--lang", "bff_noheads", "--seed", "3"],
        "selfrep_bff_noheads": ["--lang", "bff_noheads", "--seed", "4", "--eval_selfrep"],
        "selfrep_forth": ["--lang", "forth", "--seed", "5", "--eval_selfrep"],
        "fixed_shuffle": ["--lang", "bff", "--fixed_shuffle"],
        "cyclic": ["--lang", "bff_noheads", "--nopermute_programs"],
        "zero_init_initial": ["--lang", "bff", "--zero_init", "--initial_program",
                              "[[{.>]-]0000]-]>.{[[<<+-0"],
        "reset": ["--lang", "subleq", "--reset_interval", "40"],
        "mutation": ["--lang", "rsubleq4", "--mutation_prob", "0.001"],
        "pattern_2d": ["--lang", "bff_noheads", "--interaction_pattern", "pattern.txt",
                       "--seed", "8"],
        "checkpoints": ["--lang", "forthtrivial", "--checkpoint_dir", "ck"],
        "frames_1d": ["--lang", "bff", "--draw_to", "f1", "--max_epochs", "32"],
        "frames_2d": ["--lang", "forthcopy", "--draw_to_2d", "f2", "--max_epochs", "32"],
        "frames_2d_w": ["--lang", "subleq", "--draw_to_2d", "f3", "--grid_width_2d", "64",
                        "--max_epochs", "32"],
        "stop_bpb": ["--lang", "bff_noheads_4bit", "--stopping_bpb", "7.5",
                     "--max_epochs", "100000"],
    }
    cases = {}
    for name, extra in specs.items():
        d = tempfile.mkdtemp(dir=tmp)
        for sub in ("f1", "f2", "f3"):
            os.makedirs(os.path.join(d, sub))
        shutil.copyfile(os.path.join(tmp, "pattern.txt"), os.path.join(d, "pattern.txt"))
        args = base + extra
        rc, out, err = run(args, cwd=d)
        assert rc == 0, (name, err)
        files = {}
        for root, _, fs in os.walk(d):
            for fn in fs:
                rel = os.path.relpath(os.path.join(root, fn), d)
                if rel in ("log.csv", "pattern.txt"):
                    continue
                with open(os.path.join(root, fn), "rb") as f:
                    files[rel] = hashlib.sha256(f.read()).hexdigest()
        cases[name] = {"args": args, "log": open(os.path.join(d, "log.csv")).read(),
                       "files": files}
    # Resume from the epoch-64 checkpoint of the "checkpoints" case.
    d = tempfile.mkdtemp(dir=tmp)
    run(base + specs["checkpoints"], cwd=d)
    args = ["--disable_output", "--log", "log.csv", "--lang", "forthtrivial", "--load",
            "ck/0000000064.dat", "--max_epochs", "96", "--print_interval", "16",
            "--save_interval", "32"]
    rc, _, err = run(args, cwd=d)
    assert rc == 0, err
    cases["resume"] = {"args": args, "setup": "checkpoints",
                       "log": open(os.path.join(d, "log.csv")).read(), "files": {}}
    shutil.rmtree(tmp)
    dump("cli_cases.json", {
        "source": "reference bin/main built in oracle/_ref (libbrotlienc 1.0.9)",
        "runs": runs, "eval_selfrep": selfrep, "help": help_err.decode(), "errors": errors,
        "pattern_2d": "make_2d_pattern.py 32 32 (radius 2)", "cases": cases})


def gen_cli_debug():
    """Reference outputs of the debug printing: bin/main --run ... --debug
    (every step's state) and EvalParsedSelfrep with debug (every trial's
    tapes), zlib-compressed."""
    main = R.MAIN_PATH
    tmp = tempfile.mkdtemp()
    runs = {}
    for lang, texts in run_programs().items():
        rows = []
        for k, t in enumerate(texts):
            for steps in ((300, 1000) if k == 0 else (300,)):
                p = subprocess.run([main, "--lang", lang, "--run", t, "--run_steps", str(steps),
                                    "--debug"], cwd=tmp, capture_output=True)
                assert p.returncode == 0
                rows.append({"text": t, "steps": steps,
                             "stdout_zlib_b64": base64.b64encode(zlib.compress(p.stdout, 9)).decode()})
        runs[lang] = rows
    selfrep = {}
    code = ("import sys; sys.path.insert(0, {root!r}); import numpy as np; "
            "from oracle import refimpl as R; "
            "r = R.eval_parsed_selfrep({lang!r}, R.parse({lang!r}, {text!r}), {e}, {s}, debug=True); "
            "sys.stdout.flush(); sys.stderr.write(str(r))")
    for lang, texts in run_programs().items():
        rows = []
        for t in texts:
            if R.parse(lang, t).size > 64:
                continue
            p = subprocess.run([sys.executable, "-c", code.format(root=ROOT, lang=lang, text=t,
                                                                 e=3, s=11)],
                               capture_output=True)
            assert p.returncode == 0, p.stderr
            rows.append({"text": t, "epoch": 3, "seed": 11, "result": int(p.stderr.decode()),
                         "stdout_zlib_b64": base64.b64encode(zlib.compress(p.stdout, 9)).decode()})
            break
        selfrep[lang] = rows
    dump("cli_debug.json", {
        "source": "reference bin/main and EvalParsedSelfrep built in oracle/_ref (debug on)",
        "runs": runs, "eval_selfrep": selfrep})


def _grid_lists(w, h, r=2):
    return [[y * w + x for y in range(i - r, i + r + 1) for x in range(j - r, j + r + 1)
             if 0 <= y < h and 0 <= x < w and (y, x) != (i, j)]
            for i in range(h) for j in range(w)]


GENS = {"eval": gen_eval, "epochs": gen_epochs, "selfrep": gen_selfrep,
        "2d": gen_2d, "modes": gen_modes, "full": gen_full_size, "long": gen_long_run,
        "cli": gen_cli_logs, "cli_cases": gen_cli_cases, "cli_debug": gen_cli_debug}

if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=list(GENS))
    args = ap.parse_args()
    if not R.available():
        sys.exit("oracle/_ref is not built: run make -f oracle/ref.mk")
    for name in args.only:
        GENS[name]()
