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


def test_generated_site_abi_is_current():
    """include/qla_mi355x_sites.h and qla_sites_abi.inc match the variant table."""
    from qla_amd import abi_gen
    assert abi_gen.current()
    names = [v.name for v in abi_gen.variants()]
    assert len(names) == len(set(names)) == 1200
    every = [v.name for v in abi_gen.all_symbols()]
    assert len(every) == len(set(every)) == 1981
    mf = [v.name for v in abi_gen.matfuncs()]
    assert len(mf) == 321
    for s in ("QLA_F3_M_veq_expA_M", "QLA_D3_C_xeq_det_pM", "QLA_FN_M_veq_log_M", "QLA_F3_V_veq_eigenvalsH_M",
              "QLA_D3_V_veq_pM_inverse_pV", "QLA_F3_xM_eq_xM_inverse_M", "QLA_FN_M_veq_exp_M"):
        assert s in mf, s
    # QLA names of the reference's own callers and tests
    for s in ("QLA_F3_D_vpeq_spproj_M_times_pD", "QLA_D3_V_vpeq_nM_times_npV", "QLA_F3_H_veq_spproj_Ma_times_pD",
              "QLA_F3_D_xmeq_sprecon_pMa_times_pH", "QLA_D3_V_xeqm_npMa_times_nV",
              "QLA_F3_xH_eq_spproj_xM_times_D", "QLA_D3_V_peq_M_times_xV"):
        assert s in names, s


def test_header_compiles_as_c_and_cxx(tmp_path):
    inc = qlib.HEADERS[0].parent
    src = tmp_path / "use.c"
    src.write_text('#include "qla_mi355x.h"\n'
                   'void f(QLA_F3_DiracFermion *r, QLA_F3_ColorMatrix *u, QLA_F3_DiracFermion **p, int n)\n'
                   '{ QLA_F3_D_vpeq_spproj_M_times_pD(r, u, p, 0, 1, n); }\n'
                   'void g(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix **a, QLA_F_Complex *d, int *ix, int n)\n'
                   '{ QLA_F3_M_veq_expTA_pM(r, a, n); QLA_F3_C_xeq_det_M(d, r, ix, n); }\n'
                   'void h(int nc, QLA_F_Complex (*r)[4][4], QLA_F_Complex (*a)[4][4], int n)\n'
                   '{ QLA_FN_M_veq_log_M(nc, r, a, n); }\n')
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        args = [cc, std, "-fsyntax-only", "-Wall", "-Werror", "-I", str(inc)]
        if cc == "g++":
            args += ["-x", "c++"]
        subprocess.run(args + [str(src)], check=True)
