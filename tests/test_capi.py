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
    assert len(names) == len(set(names)) == 2328
    every = [v.name for v in abi_gen.all_symbols()]
    assert len(every) == len(set(every)) == 14408
    assert len(abi_gen.fieldmuls()) == 2640
    assert len(abi_gen.pmuls()) == 1440
    assert len(abi_gen.comps()) == 294
    assert len(abi_gen.sconvs()) == 174
    assert len(abi_gen.sfns()) == 502
    assert len(abi_gen.masks()) == 390
    assert len(abi_gen.ints()) == 433
    assert len(abi_gen.sums()) == 84
    assert len(abi_gen.locals_()) == 1536
    assert len(abi_gen.gammas()) == 42
    assert len(abi_gen.fills()) == 156
    mf = [v.name for v in abi_gen.matfuncs()]
    assert len(mf) == 321
    vo = [v.name for v in abi_gen.vecops()]
    assert len(vo) == 3608
    # the solver-side calls of QDP/QOP-style codes and the reference's test names
    for s in ("QLA_F3_D_veq_D", "QLA_F3_D_vpeq_r_times_D", "QLA_D3_D_veq_r_times_D_plus_D",
              "QLA_F3_D_veq_r_times_D_minus_D", "QLA_F3_D_veq_c_times_pD_plus_D", "QLA_F3_M_veq_Ma",
              "QLA_D3_P_veqm_transpose_P", "QLA_F3_D_vpeq_conj_D", "QLA_F3_D_vmeq_i_D", "QLA_F3_M_vpeq_c_times_pMa",
              "QLA_F3_xD_eq_D_minus_xD", "QLA_D3_V_eq_r_times_xV_plus_V", "QLA_F3_xM_peq_Ma",
              "QLA_F3_D_xeq_pD_plus_pD", "QLA_F3_H_veq_H_plus_H"):
        assert s in vo, s
    fl = [v.name for v in abi_gen.fills()]
    for s in ("QLA_F3_D_veq_zero", "QLA_D3_V_xeq_zero", "QLA_F3_M_vpeq_c", "QLA_D3_P_veq_p", "QLA_F3_H_xmeq_h"):
        assert s in fl, s
    for s in ("QLA_F3_M_veq_expA_M", "QLA_D3_C_xeq_det_pM", "QLA_FN_M_veq_log_M", "QLA_F3_V_veq_eigenvalsH_M",
              "QLA_D3_V_veq_pM_inverse_pV", "QLA_F3_xM_eq_xM_inverse_M", "QLA_FN_M_veq_exp_M"):
        assert s in mf, s
    # QLA names of the reference's own callers and tests
    for s in ("QLA_F3_D_vpeq_spproj_M_times_pD", "QLA_D3_V_vpeq_nM_times_npV", "QLA_F3_H_veq_spproj_Ma_times_pD",
              "QLA_F3_D_xmeq_sprecon_pMa_times_pH", "QLA_D3_V_xeqm_npMa_times_nV",
              "QLA_F3_xH_eq_spproj_xM_times_D", "QLA_D3_V_peq_M_times_xV", "QLA_F3_H_veq_spproj_D",
              "QLA_D3_D_vpeq_sprecon_pH", "QLA_F3_D_xmeq_spproj_D", "QLA_F3_xH_eq_spproj_xD",
              "QLA_F3_M_veq_M_times_Ma", "QLA_D3_D_vpeq_pMa_times_pD", "QLA_F3_H_xeq_M_times_H"):
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
                   '{ QLA_FN_M_veq_log_M(nc, r, a, n); }\n'
                   'void k(QLA_D3_DiracFermion *x, QLA_D3_DiracFermion *p, QLA_D_Real *al, QLA_D_Complex *z, int n)\n'
                   '{ QLA_D3_D_vpeq_r_times_D(x, al, p, n); QLA_D3_D_veq_c_times_D_plus_D(p + n, z, p, x, n); }\n')
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        args = [cc, std, "-fsyntax-only", "-Wall", "-Werror", "-I", str(inc)]
        if cc == "g++":
            args += ["-x", "c++"]
        subprocess.run(args + [str(src)], check=True)


def _expand(tmp_path, defines, body):
    """Preprocess `body` after including the header with `defines`; returns the
    expansion of the marked line."""
    inc = qlib.HEADERS[0].parent
    src = tmp_path / "gen.c"
    src.write_text('#include "qla_mi355x.h"\nMARK ' + body + "\n")
    out = subprocess.run(["gcc", "-std=c99", "-E", "-P", "-I", str(inc), *defines, str(src)],
                         capture_output=True, text=True, check=True).stdout
    return [ln for ln in out.splitlines() if ln.startswith("MARK")][0]


def test_generic_names_select_the_library(tmp_path):
    """QLA_Precision / QLA_Colors pick the library behind QLA's generic names
    (prototype.pl:160-260, 556-575): fully generic, colour-generic (QLA_F_,
    QLA_D_, QLA_DF_) and precision-generic (QLA_3_, QLA_N_) forms."""
    body = "QLA_V_veq_M_times_V(r, a, b, n); QLA_F_M_veq_expA_M(r, a, n); QLA_DF_r_veq_norm2_D(r, a, n);"
    e = _expand(tmp_path, [], body)
    assert "QLA_F3_V_veq_M_times_V(r, a, b, n)" in e and "QLA_F3_M_veq_expA_M" in e and "QLA_DF3_r_veq_norm2_D" in e
    e = _expand(tmp_path, ["-DQLA_Precision='D'"], "QLA_V_veq_M_times_V(r, a, b, n); QLA_3_C_veq_det_M(r, a, n);")
    assert "QLA_D3_V_veq_M_times_V(r, a, b, n)" in e and "QLA_D3_C_veq_det_M(r, a, n)" in e
    e = _expand(tmp_path, ["-DQLA_PrecisionInt=2"], "QLA_D_r_veq_norm2_V(r, a, n); QLA_F_r_veq_norm2_V(r, a, n);")
    assert "QLA_D3_r_veq_norm2_V(r, a, n)" in e and "QLA_F3_r_veq_norm2_V(r, a, n)" in e
    e = _expand(tmp_path, ["-DQLA_Precision='D'"], "QLA_r_veq_sum_R(r, a, n); QLA_c_xeq_sum_pC(r, a, ix, n);")
    assert "QLA_D_r_veq_sum_R(r, a, n)" in e and "QLA_D_c_xeq_sum_pC(r, a, ix, n)" in e
    e = _expand(tmp_path, ["-DQLA_Nc=5"], "QLA_M_veq_log_M(r, a, n); QLA_N_M_veq_exp_M(r, a, n); QLA_F_C_veq_det_M(r, a, n);")
    assert ("QLA_FN_M_veq_log_M(5, r, a, n)" in e and "QLA_FN_M_veq_exp_M(5, r, a, n)" in e
            and "QLA_FN_C_veq_det_M(5, r, a, n)" in e)
    # generic types follow the same selection; the header still compiles
    src = tmp_path / "types.c"
    src.write_text('#include "qla_mi355x.h"\n'
                   'void f(QLA_ColorMatrix *r, QLA_ColorMatrix **a, QLA_Complex *d, int n)\n'
                   '{ QLA_M_veq_expTA_pM(r, a, n); QLA_C_veq_det_M(d, r, n); }\n')
    for defs in ([], ["-DQLA_Precision='D'"]):
        subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-I", str(qlib.HEADERS[0].parent),
                        *defs, str(src)], check=True)
