"""CPU checks of the boundary and the host layer (no GPU calls):

* the HIP library loads and exports every function include/cubff_hip.h declares;
* the `cubff` Python module parses and prints programs exactly as the
  reference does (transcripts recorded from the reference CLI);
* the `main` CLI's help text and argument errors match the reference's.
"""
import base64
import json
import os
import re
import subprocess
import sys
import zlib

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
# CUBFF_HOST_BIN names another build of the host layer (the sanitizer build
# of tests/test_cpu_sanitize.py).
BIN = os.environ.get("CUBFF_HOST_BIN") or os.path.join(ROOT, "cubff_amd", "bin")
MAIN = os.path.join(BIN, "main")
RESET = "\x1b[0;38:5:35;48:5:232m"


@pytest.fixture(scope="module")
def cases():
    with open(os.path.join(GOLDEN, "cli_cases.json")) as f:
        return json.load(f)


def header_functions():
    with open(os.path.join(ROOT, "include", "cubff_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cubff_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    import ctypes

    from cubff_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = header_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert sorted(capi.EXPORTS) == names


def test_host_library_has_no_device_fallback():
    # The host layer reaches the epoch only through the HIP library: it links
    # libcubff_hip.so and none of the oracle's symbols.
    out = subprocess.run(["nm", "-D", "--undefined-only",
                          os.path.join(BIN, "libcubff_host.so")],
                         capture_output=True, text=True, check=True).stdout
    assert "cubff_run_epoch" in out
    assert "orc_" not in out


def run_py(code):
    env = dict(os.environ, PYTHONPATH=BIN)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, env=env,
                          check=True, timeout=60).stdout


def test_parse_matches_reference(cases):
    code = ("import cubff, json, sys\n"
            "rows = json.load(sys.stdin)\n"
            "print(json.dumps([bytes(cubff.GetLanguage(l).Parse(t)).hex() for l, t in rows]))\n")
    rows = [(lang, r["text"]) for lang, rs in cases["runs"].items() for r in rs]
    env = dict(os.environ, PYTHONPATH=BIN)
    out = subprocess.run([sys.executable, "-c", code], input=json.dumps(rows).encode(),
                         capture_output=True, env=env, check=True, timeout=60).stdout
    got = json.loads(out)
    want = [r["parsed"] for rs in cases["runs"].values() for r in rs]
    assert got == want


@pytest.mark.parametrize("lang", [
    "bff", "bff_noheads", "bff8", "bff8_noheads", "bff_noheads_4bit", "bff_perm",
    "bff_selfmove", "forth", "forthcopy", "forthtrivial", "forthtrivial_reset",
    "subleq", "rsubleq4"])
def test_print_program_matches_reference(cases, lang):
    # The reference's --run transcript opens with ResetColors() and the
    # zero-padded 128-byte tape as PrintProgram(128, tape, 128) prints it.
    for r in cases["runs"][lang]:
        if r["steps"] != 0:
            continue
        tape = bytes.fromhex(r["parsed"]).ljust(128, b"\0")
        out = run_py("import cubff\n"
                     f"cubff.ResetColors()\n"
                     f"cubff.GetLanguage({lang!r}).PrintProgram(128, "
                     f"cubff.VectorUint8(bytes.fromhex({tape.hex()!r})), [])\n")
        ref = zlib.decompress(base64.b64decode(r["stdout_zlib_b64"]))
        assert ref.startswith(out), (lang, r["text"])
        assert len(out) > 128


def test_cli_help_matches_reference(cases):
    p = subprocess.run([MAIN, "--help"], capture_output=True, text=True)
    assert p.returncode == 1
    # First line names argv[0].
    assert p.stderr.splitlines()[1:] == cases["help"].splitlines()[1:]


def test_cli_argument_errors_match_reference(cases, tmp_path):
    for e in cases["errors"]:
        p = subprocess.run([MAIN] + e["args"], capture_output=True, text=True, cwd=tmp_path)
        assert p.returncode == e["rc"], e["args"]
        assert p.stdout == e["stdout"], e["args"]
        if e["args"] == ["--lang", "nope"]:
            # The reference lists languages in hash-map order.
            assert sorted(p.stderr.splitlines()) == sorted(e["stderr"].splitlines())
        elif e["args"] == ["--nope"]:
            assert p.stderr.splitlines()[2:] == e["stderr"].splitlines()[2:]
        else:
            assert p.stderr == e["stderr"], e["args"]


def test_debug_fixtures_agree_with_plain_transcripts():
    # The reference's --debug transcript of a run = its plain transcript with
    # one line per step inserted after the initial program (same program,
    # same step limit).
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(golden, "cli_debug.json")) as f:
        dbg = json.load(f)["runs"]
    with open(os.path.join(golden, "cli_cases.json")) as f:
        plain = json.load(f)["runs"]
    seen = 0
    for lang, rows in dbg.items():
        for r in rows:
            for p in plain[lang]:
                if p["text"] == r["text"] and p["steps"] == r["steps"]:
                    a = zlib.decompress(base64.b64decode(r["stdout_zlib_b64"])).split(b"\n")
                    b = zlib.decompress(base64.b64decode(p["stdout_zlib_b64"])).split(b"\n")
                    assert a[0] == b[0] and a[-4:] == b[-4:], lang
                    assert len(a) > len(b) or r["steps"] == 0
                    seen += 1
    assert seen >= 13


def test_register_language_from_another_library(tmp_path):
    # A language registered by its own shared object at load time
    # (RegisterLanguage, common.h:133-134) is what GetLanguage returns, in
    # C++ and through the cubff module, beside the 13 built-in languages.
    src = os.path.join(ROOT, "tests", "plugin", "echo_language.cc")
    so = str(tmp_path / "libecho_language.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "cubff_amd", "host"), src,
                    os.path.join(BIN, "libcubff_host.so"), f"-Wl,-rpath,{BIN}", "-o", so],
                   check=True, timeout=120)
    out = run_py(
        "import ctypes\n"
        f"ctypes.CDLL({so!r}, mode=ctypes.RTLD_GLOBAL)\n"
        "import cubff\n"
        "L = cubff.GetLanguage('echo')\n"
        "print(bytes(L.Parse('abc')).decode(), flush=True)\n"
        "print(L.EvalSelfrep('xy', 3, 4, False), flush=True)\n"
        "L.RunSingleProgram('+-', 17, True)\n"
        "p = cubff.SimulationParams(); p.num_programs = 6; p.seed = 40\n"
        "seen = []\n"
        "L.RunSimulation(p, None, lambda s: seen.append((s.epoch, s.total_ops, len(s.soup))) "
        "or s.epoch >= 2)\n"
        "print(seen)\n"
        "print(bytes(cubff.GetLanguage('bff').Parse('[]')).hex())\n").decode()
    assert out.splitlines() == ["cba", "2034", "echo RunSingleProgram +- 17 1",
                                "[(0, 40, 6), (1, 41, 6), (2, 42, 6)]", "5b5d"]
