"""The CPU suite's oracle and host-layer tests under AddressSanitizer and
UndefinedBehaviorSanitizer, and the debug check for pairings that put one
program in two slots.

The oracle (oracle/cubff_oracle.c, with -DORC_DEBUG_CHECKS) and the host
layer (cubff_amd/host: RunSimulation's host side, the presenters, the CLI,
the Python module) are built with -fsanitize=address,undefined into a
temporary directory (`make -C oracle san`, `make -C cubff_amd/host san`) and
the CPU tests that exercise them run against those builds in a child
process with the sanitizer runtime preloaded; any report fails the run.

The race the reference leaves to its callers (common_language.h:165-192
assumes the epoch's pairing `s` is a permutation) is what the debug build
checks at every epoch: a program in two slots is written back by both.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    cc = shutil.which("gcc")
    if not cc:
        return None
    p = subprocess.run([cc, "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_builds(tmp_path_factory):
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not available")
    if not os.path.exists(os.path.join(ROOT, "cubff_amd", "lib", "libcubff_hip.so")):
        pytest.skip("the HIP library is not built (the host layer links it)")
    d = tmp_path_factory.mktemp("san")
    oracle_lib = str(d / "libcubff_oracle_san.so")
    host_bin = str(d / "bin")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san",
                    f"SANLIB={oracle_lib}"], check=True, timeout=300, capture_output=True)
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "cubff_amd", "host"), "san",
                    f"SAN={host_bin}"], check=True, timeout=600, capture_output=True)
    return asan, oracle_lib, host_bin


def test_cpu_suite_under_sanitizers(san_builds):
    asan, oracle_lib, host_bin = san_builds
    env = dict(os.environ,
               LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               CUBFF_ORACLE_LIB=oracle_lib,
               CUBFF_HOST_BIN=host_bin,
               OMP_NUM_THREADS="4")
    tests = [os.path.join(ROOT, "tests", t) for t in
             ("test_cpu_oracle.py", "test_cpu_host.py", "test_cpu_trace.py")]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    print(p.stdout[-3000:], p.stderr[-3000:])
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    assert "passed" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr
    assert "runtime error:" not in p.stderr


def test_debug_check_flags_pairings_that_repeat_programs(san_builds):
    # Fixed pairing on a soup whose size is not a power of two repeats
    # programs (main.cc:202 rejects it; the library's other entry points
    # accept it, as the reference's Python module does); Fisher-Yates,
    # cyclic and fixed pairing at powers of two are permutations.
    asan, oracle_lib, _ = san_builds
    code = (
        "import numpy as np\n"
        "from oracle import oracle as O\n"
        "assert O.pairing_repeats(O.pairing_fixed(6, 1), 6) > 0\n"
        "assert O.pairing_repeats(O.pairing_shuffle(4096, 3, 5), 4096) == 0\n"
        "assert O.pairing_repeats(O.pairing_cyclic(4096, 7), 4096) == 0\n"
        "assert O.pairing_repeats(O.pairing_fixed(4096, 9), 4096) == 0\n"
        "soup = O.init_programs(6, O.epoch_seed(1, 0))\n"
        "O.mutate_and_run('bff', soup, O.pairing_fixed(6, 1), 3, O.epoch_seed(1, 1), 1 << 18)\n"
        "print('seen', O.debug_repeats_seen())\n")
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               CUBFF_ORACLE_LIB=oracle_lib)
    p = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert int(p.stdout.split("seen")[1]) > 0
    assert "puts" in p.stderr and "more than one slot" in p.stderr


def test_pairing_repeats_closed_forms(oracle):
    # The check itself (normal build): a permutation has none; each program
    # beyond its first slot counts once.
    from oracle import oracle as O
    assert O.pairing_repeats(np.arange(8, dtype=np.uint32), 8) == 0
    assert O.pairing_repeats(np.array([0, 1, 0, 2, 3, 1], dtype=np.uint32), 6) == 2
    # One program in both halves of one slot (a 2-D pattern line "a a") is
    # not a repeat across slots.
    assert O.pairing_repeats(np.array([4, 4, 0, 1, 2, 3], dtype=np.uint32), 6) == 0
