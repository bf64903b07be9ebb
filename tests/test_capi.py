"""The C-ABI library loads on a machine without a GPU and exports every
symbol the public header declares (no compute calls here)."""
import subprocess

import qla_amd
from qla_amd import lib as qlib


def test_library_exports_header_symbols():
    lib = qla_amd.load()
    syms = qla_amd.header_symbols()
    assert len(syms) >= 28
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(qla_amd.header_symbols()) == set(qlib._SIGS)


def test_dynamic_symbol_table():
    out = subprocess.run(["nm", "-D", "--defined-only", str(qla_amd.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for s in qla_amd.header_symbols():
        assert s in exported, s


def test_version_and_error_state():
    assert "gfx950" in qla_amd.version()
    assert qla_amd.load().QLA_MI355X_last_error() == 0
