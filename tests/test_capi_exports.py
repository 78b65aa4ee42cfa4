"""The engine library loads on a machine without a GPU and exports every
entry point declared in include/mpm/*.h (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from mpm_mi355x import capi


def declared():
    inc = os.path.dirname(capi.HEADER_PATH)
    names = set()
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_the_reference_pass_surface():
    names = declared()
    for n in ("mpm_bin", "mpm_p2g", "mpm_solve_grid", "mpm_g2p", "mpm_step", "mpm_maksf_batch",
              "mpm_engine_create", "mpm_set_bc", "mpm_download_bins", "mpm_mlp_forward", "mpm_mlp_gradient",
              "mpm_mlp_train", "mpm_surrogate_fit", "mpm_surrogate_enable"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("engine library not built")
    lib = C.CDLL(capi.LIB_PATH)
    for n in declared():
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mpm_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_binding_signatures_cover_the_header():
    assert set(declared()) == set(capi.SIGNATURES)


def test_version_string_without_gpu():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("engine library not built")
    assert capi.load().mpm_version().decode().startswith("mpm-mi355x")
