"""The C ABI boundary: the built gfx950 library loads on a host without a GPU
and exports every entry point include/pscf_hip.h declares, and the ctypes
signature table of pscf_amd._lib covers exactly those entry points."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pscf_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(pscf_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) > 40
    for must in ("pscf_density", "pscf_scf_stress", "pscf_step_exp", "pscf_fft_r2c",
                 "pscf_basis_to_kgrid", "pscf_kgrid_to_basis", "pscf_density_adjoint_dev",
                 "pscf_batch_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pscf_amd._lib import DEFAULT_LIB
    if not os.path.exists(DEFAULT_LIB):
        pytest.skip("libpscf_hip.so not built")
    dll = ctypes.CDLL(DEFAULT_LIB)
    missing = [n for n in _declared() if not hasattr(dll, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from pscf_amd._lib import SIGNATURES
    declared = set(_declared())
    special = {"pscf_last_error", "pscf_am_last_error"}      # restype char*, set up by Lib
    assert set(SIGNATURES) | special == declared, (
        sorted(declared - set(SIGNATURES) - special), sorted(set(SIGNATURES) - declared))
