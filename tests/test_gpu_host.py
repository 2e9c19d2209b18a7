"""GPU checks of the host layer against recorded reference runs.

* `main` (the CLI) writes the same CSV logs, checkpoints and frames as the
  reference CLI for every mode flag, and the same --run transcripts;
* the test.sh workload (13 languages, 131072 programs, 256 epochs, seed
  10248) logs exactly what the reference logs with this image's libbrotli
  (1.0.9), and the same zero-order entropy as the reference's own
  testdata/ files, which were written with another libbrotli;
* the `cubff` Python module's RunSimulation reaches the reference's soup
  hashes at full size, and EvalParsedSelfrep scores match.
"""
import base64
import hashlib
import json
import os
import subprocess
import sys
import zlib

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "cubff_amd", "bin")
MAIN = os.path.join(BIN, "main")
LANGS = ["bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit", "bff_perm",
         "bff_selfmove", "forth", "forthcopy", "forthtrivial", "forthtrivial_reset",
         "subleq", "rsubleq4"]


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def run_main(args, cwd, timeout=120):
    print("main", " ".join(args), flush=True)
    p = subprocess.run([MAIN] + args, cwd=cwd, capture_output=True, timeout=timeout)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    return p.stdout


@pytest.mark.parametrize("lang", LANGS)
def test_cli_log_matches_reference(lang, tmp_path):
    # test.sh:30 settings.
    run_main(["--lang", lang, "--max_epochs", "256", "--disable_output", "--log", "log.txt",
              "--seed", "10248"], tmp_path)
    got = (tmp_path / "log.txt").read_text()
    assert got == load("cli_logs_brotli109.json")["logs"][lang]
    # The reference's own testdata: same soups, hence the same H0 =
    # higher_entropy + brotli_bpb (to print precision), different brotli sizes.
    with open(os.path.join(GOLDEN, "reference_testdata", lang + ".txt")) as f:
        ref = f.read().splitlines()[1:]
    for a, b in zip(got.splitlines()[1:], ref, strict=True):
        ea, sa, na, ha = a.split(",")
        eb, sb, nb, hb = b.split(",")
        assert (ea, na) == (eb, nb)
        bits = 8.0 / (int(na) * 64)
        h0a = float(ha) + int(sa) * bits
        h0b = float(hb) + int(sb) * bits
        assert abs(h0a - h0b) < 2e-6, (lang, ea, h0a, h0b)


@pytest.mark.parametrize("lang", LANGS)
def test_cli_run_transcripts_match_reference(lang, tmp_path):
    for r in load("cli_cases.json")["runs"][lang]:
        out = run_main(["--lang", lang, "--run", r["text"], "--run_steps", str(r["steps"])],
                       tmp_path)
        ref = zlib.decompress(base64.b64decode(r["stdout_zlib_b64"]))
        assert out == ref, (lang, r["text"], r["steps"])


CASES = ["shuffle_bff_noheads", "selfrep_bff_noheads", "selfrep_forth", "fixed_shuffle",
         "cyclic", "zero_init_initial", "reset", "mutation", "pattern_2d", "checkpoints",
         "frames_1d", "frames_2d", "frames_2d_w", "stop_bpb", "resume"]


def grid_lines(w, h, r=2):
    """make_2d_pattern.py:17-29 output for a w x h grid."""
    out = []
    for i in range(h):
        for j in range(w):
            for y in range(i - r, i + r + 1):
                for x in range(j - r, j + r + 1):
                    if 0 <= y < h and 0 <= x < w and (y, x) != (i, j):
                        out.append(f"{i * w + j} {y * w + x}")
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("name", CASES)
def test_cli_cases_match_reference(name, tmp_path):
    c = load("cli_cases.json")["cases"][name]
    for sub in ("f1", "f2", "f3"):
        (tmp_path / sub).mkdir()
    (tmp_path / "pattern.txt").write_text(grid_lines(32, 32))
    if "setup" in c:
        run_main(load("cli_cases.json")["cases"][c["setup"]]["args"], tmp_path)
    run_main(c["args"], tmp_path, timeout=300)
    assert (tmp_path / "log.csv").read_text() == c["log"]
    for rel, digest in c["files"].items():
        with open(tmp_path / rel, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == digest, rel
    made = {str(p.relative_to(tmp_path)) for p in tmp_path.rglob("*") if p.is_file()}
    assert made - {"log.csv", "pattern.txt"} >= set(c["files"])


def py_module():
    if BIN not in sys.path:
        sys.path.insert(0, BIN)
    import cubff
    return cubff


def test_python_module_full_size_soups():
    cubff = py_module()
    g = load("full_size.json")
    lang = cubff.GetLanguage("bff_noheads")
    for seed, rows in g["seeds"].items():
        print("RunSimulation seed", seed, flush=True)
        params = cubff.SimulationParams()
        params.num_programs = 131072
        params.seed = int(seed)
        params.callback_interval = 64
        seen = []

        def cb(state):
            seen.append((state.epoch, hashlib.sha256(memoryview(state.soup)).hexdigest(),
                         state.total_ops, len(state.shuffle_idx)))
            return state.epoch >= 257

        lang.RunSimulation(params, None, cb)
        assert [s[0] for s in seen] == [r["epoch"] for r in rows]
        assert [s[1] for s in seen] == [r["soup"] for r in rows]
        assert [s[2] for s in seen] == [r["total_ops"] for r in rows]
        assert all(s[3] == 131072 for s in seen)


def test_python_module_long_run():
    # The 1000-epoch bff_noheads run at 2^17 (seed 0, callbacks every 64
    # epochs, so 1025 epochs as the reference CLI runs it): soup and pairing
    # hashes and cumulative ops at every callback equal the reference's.
    cubff = py_module()
    g = load("bff_noheads_2p17_seed0_1025_epochs.json")
    p = g["params"]
    params = cubff.SimulationParams()
    params.num_programs = p["num_programs"]
    params.seed = p["seed"]
    params.callback_interval = p["callback_interval"]
    seen = []

    def cb(state):
        seen.append({"epoch": state.epoch,
                     "soup": hashlib.sha256(memoryview(state.soup)).hexdigest(),
                     "shuffle": hashlib.sha256(memoryview(state.shuffle_idx)).hexdigest(),
                     "total_ops": state.total_ops})
        if state.epoch % 256 == 1:
            print("long run at epoch", state.epoch, flush=True)
        return state.epoch >= p["epochs"]

    print("RunSimulation bff_noheads 2^17 seed 0 to epoch 1025", flush=True)
    cubff.GetLanguage(p["lang"]).RunSimulation(params, None, cb)
    assert seen == g["rows"]


@pytest.mark.parametrize("lang", LANGS)
def test_cli_debug_transcripts_match_reference(lang, tmp_path):
    # --run ... --debug: the state before every step, printed the way each
    # language's Evaluate prints it, from step records made on the GPU.
    for r in load("cli_debug.json")["runs"][lang]:
        out = run_main(["--lang", lang, "--run", r["text"], "--run_steps", str(r["steps"]),
                        "--debug"], tmp_path)
        ref = zlib.decompress(base64.b64decode(r["stdout_zlib_b64"]))
        assert out == ref, (lang, r["text"], r["steps"])


def test_python_module_eval_selfrep_debug(tmp_path):
    # EvalSelfrep with debug prints every trial's tapes before and after each
    # generation, as CheckSelfRep does; the score is unchanged.
    script = tmp_path / "sr.py"
    script.write_text(
        "import sys, json, base64\n"
        f"sys.path.insert(0, {BIN!r})\n"
        "import cubff\n"
        "lang, text, e, s = sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4])\n"
        "r = cubff.GetLanguage(lang).EvalSelfrep(text, e, s, True)\n"
        "sys.stdout.flush()\n"
        "sys.stderr.write('RESULT %d' % r)\n")
    for lang, rows in load("cli_debug.json")["eval_selfrep"].items():
        for r in rows:
            print("selfrep debug", lang, r["text"], flush=True)
            p = subprocess.run([sys.executable, str(script), lang, r["text"], str(r["epoch"]),
                                str(r["seed"])], capture_output=True, timeout=120)
            assert p.returncode == 0, p.stderr.decode()[-2000:]
            assert p.stderr.decode().endswith("RESULT %d" % r["result"]), lang
            ref = zlib.decompress(base64.b64decode(r["stdout_zlib_b64"]))
            assert p.stdout == ref, lang


def test_python_module_eval_selfrep():
    cubff = py_module()
    for lang, rows in load("cli_cases.json")["eval_selfrep"].items():
        L = cubff.GetLanguage(lang)
        for r in rows:
            if "text" in r:
                got = L.EvalSelfrep(r["text"], r["epoch"], r["seed"], False)
            else:
                got = L.EvalParsedSelfrep(cubff.VectorUint8(bytes.fromhex(r["parsed"])),
                                          r["epoch"], r["seed"], False)
            assert got == r["result"], (lang, r)
