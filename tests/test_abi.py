"""The C ABI: both libraries load and export exactly the functions that
include/reseact_hip.h declares (no compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from reseact_esm_amd.host import ABI, HIP_LIB_PATH

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "reseact_hip.h")


def declared() -> set[str]:
    text = open(HEADER).read()
    return set(re.findall(r"RS_API\s+(?:int|const char\*)\s+(rs_\w+)\(", text))


def exported(path: str) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_header_lists_binding_table():
    d = declared()
    assert len(d) >= 30
    assert d == set(ABI), f"header vs binding table: {sorted(d ^ set(ABI))}"


@pytest.mark.skipif(not os.path.exists(HIP_LIB_PATH), reason="HIP library not built")
def test_hip_library_exports_exactly_the_header():
    ex = exported(HIP_LIB_PATH)
    assert ex == declared(), f"difference: {sorted(ex ^ declared())}"
    lib = ctypes.CDLL(HIP_LIB_PATH)
    lib.rs_backend.restype = ctypes.c_char_p
    assert lib.rs_backend().decode() == "hip-gfx950"


def test_oracle_exports_header_plus_extras(oracle):
    ex = exported(oracle.path)
    assert declared() <= ex
    assert {s for s in ex - declared()} == {s for s in ex if s.startswith("rs_oracle_")}
    assert oracle.backend == "cpu-oracle"


@pytest.mark.skipif(not os.path.exists(HIP_LIB_PATH), reason="HIP library not built")
def test_hip_library_code_object_is_gfx950():
    blob = open(HIP_LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"gfx90a" not in blob and b"gfx942" not in blob


def test_errors_are_reported(oracle):
    ctx = oracle.context()
    with pytest.raises(Exception) as e:
        ctx._chk(ctx.lib.rs_set_grid(ctx.h, 0, 1, 1, 1, 1))
    assert "RS_E_ARG" in str(e.value)
    assert "bad grid" in ctx.lib.rs_last_error(ctx.h).decode()
