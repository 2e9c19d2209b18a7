"""CPU checks of the BFF interpreters (cubff_amd/csrc/bff_fast.h,
cubff_amd/csrc/bff_lean.h and cubff_amd/csrc/bff_wave.h): their templates
compiled for the host (tests/host_interp/bff_fast_check.hip,
bff_lean_check.hip, bff_wave_check.hip) against the
oracle's Evaluate on random tapes, step limits and command budgets, for the
four languages of the standard jump logic.  The oracle is itself pinned by
the reference build's vectors (tests/test_cpu_oracle.py)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory, oracle):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("interp") / "bff_fast_check")
    lib = os.path.join(ROOT, "oracle", "libcubff_oracle.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--cuda-host-only", "-x", "hip",
                    os.path.join(ROOT, "tests", "host_interp", "bff_fast_check.hip"),
                    "-x", "none", lib, f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", exe],
                   check=True, timeout=300, capture_output=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_bff_fast_matches_oracle(checker, seed):
    p = subprocess.run([checker, "60000", str(seed)], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("tapes ok") == 8


@pytest.fixture(scope="module")
def lean_checker(tmp_path_factory, oracle):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("interp") / "bff_lean_check")
    lib = os.path.join(ROOT, "oracle", "libcubff_oracle.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--cuda-host-only", "-x", "hip",
                    os.path.join(ROOT, "tests", "host_interp", "bff_lean_check.hip"),
                    "-x", "none", lib, f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", exe],
                   check=True, timeout=300, capture_output=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_bff_lean_matches_oracle(lean_checker, seed):
    """The word-table interpreter (cubff_amd/csrc/bff_lean.h) on the host:
    whole runs and runs resumed in slices, at random step limits and
    budgets, equal the oracle's Evaluate, and the tables it keeps up to
    date through every write equal a fresh build of the final tape."""
    p = subprocess.run([lean_checker, "60000", str(seed)], capture_output=True, text=True,
                       timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("tapes ok") == 4


@pytest.fixture(scope="module")
def wave_checker(tmp_path_factory, oracle):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("interp") / "bff_wave_check")
    lib = os.path.join(ROOT, "oracle", "libcubff_oracle.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--cuda-host-only", "-x", "hip",
                    os.path.join(ROOT, "tests", "host_interp", "bff_wave_check.hip"),
                    "-x", "none", lib, f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", exe],
                   check=True, timeout=300, capture_output=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_bff_wave_matches_oracle(wave_checker, seed):
    """The resolver's wave-wide interpreter (cubff_amd/csrc/bff_wave.h) on
    the host, the wavefront's registers as an array: runs from the start
    and runs taken over from a word-table run stopped by its budget equal
    the oracle's Evaluate."""
    p = subprocess.run([wave_checker, "60000", str(seed)], capture_output=True, text=True,
                       timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("tapes ok") == 4
