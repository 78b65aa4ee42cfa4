"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

`make -C tests/cpp sanitize` builds, with -fsanitize=address,undefined and no
recovery:
- ref_shim_test_san: the reference-signature shim (shim/ref_shim.cpp) driven
  by the caller written like the reference's passes, with the FP64 oracle
  compiled in as its checker;
- oracle_san: the oracle's step over every constitutive model, transfer
  scheme and shape-function order, with shifting, F-bar, two materials,
  traction and boundary lists, and the network oracle;
- case_io_san: the case-file and restart-file readers and writers
  (csrc/case_io.cpp) on the beam case, on copies cut short or with a byte
  replaced, and on restart round trips.

A sanitizer report aborts the program, so each run must end with status 0
and print OK. The GPU kernels are not built here (no GPU sanitizer is used).
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpp", "_build")
sys.path.insert(0, os.path.join(HERE, "golden"))


def _have_sanitizers(tmp_path):
    if not (shutil.which("gcc") and shutil.which("g++") and shutil.which("make")):
        return False
    src = tmp_path / "t.c"
    src.write_text("int main(void){return 0;}\n")
    r = subprocess.run(["gcc", "-fsanitize=address,undefined", str(src), "-o", str(tmp_path / "t")],
                       capture_output=True)
    return r.returncode == 0


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not _have_sanitizers(tmp_path_factory.mktemp("probe")):
        pytest.skip("gcc without AddressSanitizer / UndefinedBehaviorSanitizer")
    r = subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return BUILD


def _run(args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and out.rstrip().endswith("OK"), out[-3000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-3000:]
    return out


def test_shim_caller_under_sanitizers(built):
    _run([os.path.join(built, "ref_shim_test_san")])


def test_oracle_step_under_sanitizers(built):
    out = _run([os.path.join(built, "oracle_san")])
    assert out.count("finite") >= 24


def test_case_and_restart_files_under_sanitizers(built, tmp_path):
    import beam_case
    beam_case.make(str(tmp_path))
    out = _run([os.path.join(built, "case_io_san"), beam_case.BEAM_INPUT, str(tmp_path / "griddata.txt"),
                str(tmp_path / "spdata.txt"), str(tmp_path)])
    assert "written and read back" in out and "damaged 60 times" in out
