"""The site-operation part of the C ABI, generated from one variant table.

QLA's generator emits its Wilson-hop building blocks for every combination of
assignment operator, adjoint, pointer operand and gang index
(lib/perl/build_qla.pl:1087-1250 and 1693-1725, looping over
operatortypes.pl:34 ``@eqop_all`` and the indexing list of
indirection.pl:63-120; names and argument order from prototype.pl:450-688).
This module enumerates the same combinations for QLA_F3 and QLA_D3:

=============================  ==========================================  =======
family                         symbol pattern                              kind
=============================  ==========================================  =======
matrix times vector            V_{v,x}<eqop>_{M,Ma,pM,pMa}_times_{V,pV}      0
projection, then matrix        H_{v,x}<eqop>_spproj_<A>_times_{D,pD}         1
matrix, then reconstruction    D_{v,x}<eqop>_sprecon_<A>_times_{H,pH}        2
project-multiply-reconstruct   D_{v,x}<eqop>_spproj_<A>_times_{D,pD}         3
multi-source matrix times vec  V_{v,x}<eqop>_n<A>_times_n{V,pV}  (+ int nd)  multi
spin projection                H_{v,x}<eqop>_spproj_{D,pD}                   5
spin reconstruction            D_{v,x}<eqop>_sprecon_{H,pH}                  6
projection with reconstruction D_{v,x}<eqop>_spproj_{D,pD}                   7
matrix times H, D, M, M^dag    {H,D,M}_{v,x}<eqop>_<A>_times_[p]{H,D,M,Ma}  8-11
=============================  ==========================================  =======

with <eqop> in eq, peq, meq, eqm; and the mixed-precision field layer
(build_qla.pl:530-585 conversions, :2089-2360 global reductions) for the
types V, H, D, M, P:

=============================  ==========================================  =======
family                         symbol pattern                              libs
=============================  ==========================================  =======
global squared norm            r_{v,x}eq_norm2_{T,pT}                      F3 D3 DF3
global inner product           c_{v,x}eq_[p]T_dot_[p]T                     F3 D3 DF3
real part of inner product     r_{v,x}eq_re_[p]T_dot_[p]T                  F3 D3 DF3
precision conversion           T_{v,x}eq_[p]T  (float -> double, back)     DF3 FD3
=============================  ==========================================  =======

and the colour-matrix functions (build_qla.pl:1002-1066, 1663-1690) for
QLA_F3, QLA_D3 and QLA_FN (leading ``int nc``, nc = 3..8):

=============================  ==========================================
family                         symbol pattern
=============================  ==========================================
determinant                    C_{v,x}eq_det_[p]M
Hermitian eigenvalues (nc=3)   V_{v,x}eq_eigenvalsH_[p]M
matrix function                M_{v,x}eq_<f>_[p]M, f in inverse, sqrt,
                               sqrtPH, invsqrt, invsqrtPH, exp, expA,
                               expTA, log
inverse times vector / matrix  {V,M}_{v,x}eq_[p]M_inverse_[p]{V,M}
=============================  ==========================================

each also in the per-operand index forms (``xM_eq_expA_xM`` ...).

It writes

* ``include/qla_mi355x_{sites,fields,matfuncs}.h`` -- the prototypes,
* ``include/qla_mi355x_generic.h`` -- QLA's generic names for them, and
* ``qla_amd/csrc/qla_{sites,fields,matfuncs}_abi.inc`` -- their definitions,
  one call each into the generic entry helpers of ``qla_capi.hip``.

Both files are committed; ``python -m qla_amd.abi_gen`` rewrites them and a CPU
test checks that they match this table.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass
from pathlib import Path

HERE = Path(__file__).resolve().parent
HEADER_OUT = HERE.parent / "include" / "qla_mi355x_sites.h"
INC_OUT = HERE / "csrc" / "qla_sites_abi.inc"
FIELDS_HEADER_OUT = HERE.parent / "include" / "qla_mi355x_fields.h"
FIELDS_INC_OUT = HERE / "csrc" / "qla_fields_abi.inc"
MATFUNCS_HEADER_OUT = HERE.parent / "include" / "qla_mi355x_matfuncs.h"
MATFUNCS_INC_OUT = HERE / "csrc" / "qla_matfuncs_abi.inc"
GENERIC_HEADER_OUT = HERE.parent / "include" / "qla_mi355x_generic.h"
VECOPS_HEADER_OUT = HERE.parent / "include" / "qla_mi355x_vecops.h"
VECOPS_INC_OUT = HERE / "csrc" / "qla_vecops_abi.inc"

PRECS = ("F3", "D3")
EQOPS = ("eq", "peq", "meq", "eqm")          # codes 0..3, as the kernels use them
TYPE_NAMES = {"V": "ColorVector", "H": "HalfFermion", "D": "DiracFermion", "M": "ColorMatrix",
              "P": "DiracPropagator"}
# complex numbers per record (include/qla_types.h:29-42)
NCOMP = {"V": 3, "H": 6, "D": 12, "M": 9, "P": 144, "C": 1, "R": 0}   # R: one real (the kernels' code 0)
# the real and complex fields live in the colourless libraries QLA_F / QLA_D
SCALAR_TYPES = {"R": "Real", "C": "Complex"}


def lib_of(prec: str, t: str) -> str:
    """library prefix of a record type: F3 / D3, or F / D for R and C"""
    return prec[0] if t in SCALAR_TYPES else prec


def rec_type(prec: str, t: str) -> str:
    """C type of a record of type t in library prec"""
    if t == "I":
        return "QLA_Int"
    if t in SCALAR_TYPES:
        return f"QLA_{prec[0]}_{SCALAR_TYPES[t]}"
    return f"QLA_{prec}_{TYPE_NAMES[t]}"
FIELD_TYPES = ("V", "H", "D", "M", "P")
# (dest, src1, src2) index flags of the per-operand index forms of a binary op
OWN_INDEX_FORMS = [(r, a, b) for r in (False, True) for a in (False, True) for b in (False, True) if r or a or b]
KINDS = {"matvec": 0, "spproj": 1, "sprecon": 2, "wilson": 3, "multi": 4, "proj": 5, "recon": 6, "projrecon": 7,
         "mat_h": 8, "mat_d": 9, "mat_m": 10, "mat_ma": 11}
# the matrix-free spin operations (one source, `a` in the QLA name)
UNARY_KINDS = ("proj", "recon", "projrecon")
# (dest record, second-source record, qualifier in the name)
FAMILY = {
    "matvec": ("V", "V", ""),
    "spproj": ("H", "D", "spproj_"),
    "sprecon": ("D", "H", "sprecon_"),
    "wilson": ("D", "D", "spproj_"),
    "multi": ("V", "V", ""),
    "proj": ("H", "D", "spproj_"),
    "recon": ("D", "H", "sprecon_"),
    "projrecon": ("D", "D", "spproj_"),
    "mat_h": ("H", "H", ""),
    "mat_d": ("D", "D", ""),
    "mat_m": ("M", "M", ""),
    "mat_ma": ("M", "M", ""),        # second factor adjoint ("Ma")
}


@dataclass(frozen=True)
class Variant:
    prec: str       # "F3" / "D3"
    kind: str       # key of KINDS
    eqop: int       # 0 eq, 1 peq, 2 meq, 3 eqm
    adj: bool       # A = M^dagger ("Ma")
    xf: bool        # gang-indexed x-form
    pa: bool        # A is a pointer array ("pM")
    pb: bool        # B is a pointer array ("pV", "pD", "pH")
    ir: bool = False    # per-operand index forms of QLA_INDEX_ALL (indirection_all.pl:66-84):
    ia: bool = False    # r[indexr[i]], a[indexa[i]], b[indexb[i]] ("xH", "xM", "xD")
    ib: bool = False

    @property
    def own_index(self) -> bool:
        return self.ir or self.ia or self.ib

    @property
    def multi(self) -> bool:
        return self.kind == "multi"

    @property
    def dest(self) -> str:
        return FAMILY[self.kind][0]

    @property
    def src2(self) -> str:
        return FAMILY[self.kind][1]

    @property
    def spin(self) -> bool:
        return self.kind in ("spproj", "sprecon", "wilson") + UNARY_KINDS

    @property
    def unary(self) -> bool:
        return self.kind in UNARY_KINDS

    @property
    def name(self) -> str:
        form = "" if self.own_index else ("x" if self.xf else "v")
        if self.unary:
            src = f"{'p' if self.pa else ''}{'x' if self.ia else ''}{self.src2}"
            r = f"{'x' if self.ir else ''}{self.dest}"
            return f"QLA_{self.prec}_{r}_{form}{EQOPS[self.eqop]}_{FAMILY[self.kind][2]}{src}"
        n = "n" if self.multi else ""
        a = f"{n}{'p' if self.pa else ''}{'x' if self.ia else ''}M{'a' if self.adj else ''}"
        b = f"{n}{'p' if self.pb else ''}{'x' if self.ib else ''}{self.src2}{'a' if self.kind == 'mat_ma' else ''}"
        r = f"{'x' if self.ir else ''}{self.dest}"
        return f"QLA_{self.prec}_{r}_{form}{EQOPS[self.eqop]}_{FAMILY[self.kind][2]}{a}_times_{b}"

    def params(self) -> list[tuple[str, str]]:
        """(kind, C declaration) of every parameter; kind 'p' pointer or 'i' int."""
        t = lambda s: f"QLA_{self.prec}_{TYPE_NAMES[s]}"                  # noqa: E731
        stars = lambda ptr: "*" * ((1 if self.multi else 0) + (1 if ptr else 0))  # noqa: E731
        out = [("p", f"{t(self.dest)} *QLA_MI355X_RESTRICT r")]
        if self.unary:
            if self.ir:
                out.append(("p", "int *indexr"))
            out.append(("p", f"{t(self.src2)} *QLA_MI355X_RESTRICT {stars(self.pa)}a"))
            if self.ia:
                out.append(("p", "int *indexa"))
            out += [("i", "int mu"), ("i", "int sign")]
            if self.xf:
                out.append(("p", "int *index"))
            out.append(("i", "int n"))
            return out
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{t('M')} *QLA_MI355X_RESTRICT {stars(self.pa)}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out.append(("p", f"{t(self.src2)} *QLA_MI355X_RESTRICT {stars(self.pb)}b"))
        if self.ib:
            out.append(("p", "int *indexb"))
        if self.spin:
            out += [("i", "int mu"), ("i", "int sign")]
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        if self.multi:
            out.append(("i", "int nd"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())


@dataclass(frozen=True)
class RedVariant:
    """A global reduction: norm2 of a, or conj(a).b summed (dot / re dot)."""
    prec: str       # "F3", "D3" or "DF3" (float fields, double result)
    kind: str       # "norm2", "dot", "redot"
    t: str          # record type letter
    xf: bool
    pa: bool
    pb: bool        # dot / redot only
    ia: bool = False    # per-operand index forms of the binary reductions
    ib: bool = False    # (QLA_INDEX_ALL): a[indexa[i]], b[indexb[i]]

    @property
    def name(self) -> str:
        form = "" if (self.ia or self.ib) else ("x" if self.xf else "v")
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + self.t
        if self.kind == "norm2":
            return f"QLA_{self.prec}_r_{form}eq_norm2_{a}"
        b = ("p" if self.pb else "") + ("x" if self.ib else "") + self.t
        lead = "c" if self.kind == "dot" else "r"
        qual = "re_" if self.kind == "redot" else ""
        return f"QLA_{self.prec}_{lead}_{form}eq_{qual}{a}_dot_{b}"

    def params(self) -> list[tuple[str, str]]:
        rp = "F" if self.prec == "F3" else "D"
        fp = "D3" if self.prec == "D3" else "F3"
        rt = f"QLA_{rp}_Complex" if self.kind == "dot" else f"QLA_{rp}_Real"
        ft = f"QLA_{fp}_{TYPE_NAMES[self.t]}"
        out = [("p", f"{rt} *QLA_MI355X_RESTRICT r"),
               ("p", f"{ft} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a")]
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.kind != "norm2":
            out.append(("p", f"{ft} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
        if self.ib:
            out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        # result kinds of the device combine: 0 double, 1 float, 2 double complex,
        # 3 float complex, 4 double real part, 5 float real part
        out = {("norm2", True): 1, ("norm2", False): 0, ("dot", True): 3, ("dot", False): 2,
               ("redot", True): 5, ("redot", False): 4}[(self.kind, self.prec == "F3")]
        nbytes = {0: 8, 1: 4, 2: 16, 3: 8, 4: 8, 5: 4}[out]
        b = "(const void *)b" if self.kind != "norm2" else "nullptr"
        idx = "index" if self.xf else "nullptr"
        ia = "indexa" if self.ia else "nullptr"
        ib = "indexb" if self.ib else "nullptr"
        return (f"reduce_entry({prec}, {int(self.kind != 'norm2')}, {NCOMP[self.t]}, {int(self.pa)}, "
                f"{int(self.pb)}, r, (const void *)a, {b}, {idx}, {ia}, {ib}, n, {out}, {nbytes}, "
                f"\"{self.name}\");")


@dataclass(frozen=True)
class ConvVariant:
    """Precision conversion of a field: DF3 float -> double, FD3 double -> float."""
    lib: str        # "DF3" or "FD3"
    t: str
    xf: bool
    pa: bool
    ir: bool = False    # per-operand index forms (QLA_INDEX_ALL): r[indexr[i]],
    ia: bool = False    # a[indexa[i]]

    @property
    def name(self) -> str:
        form = "" if (self.ir or self.ia) else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + self.t
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + self.t
        return f"QLA_{self.lib}_{r}_{form}eq_{a}"

    def params(self) -> list[tuple[str, str]]:
        dst, src = ("D3", "F3") if self.lib == "DF3" else ("F3", "D3")
        out = [("p", f"QLA_{dst}_{TYPE_NAMES[self.t]} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"QLA_{src}_{TYPE_NAMES[self.t]} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        idx = "index" if self.xf else "nullptr"
        ir = "indexr" if self.ir else "nullptr"
        ia = "indexa" if self.ia else "nullptr"
        return (f"convert_entry({int(self.lib == 'FD3')}, {NCOMP[self.t]}, {int(self.pa)}, r, (const void *)a, "
                f"{idx}, {ir}, {ia}, n, \"{self.name}\");")


# colour-matrix functions: name -> (kernel code, result type, second operand type)
MATFUNCS = {
    "exp": (0, "M", None), "sqrt": (1, "M", None), "invsqrt": (2, "M", None), "det": (3, "C", None),
    "inverse": (4, "M", None), "inverse_V": (5, "V", "V"), "inverse_M": (6, "M", "M"),
    "expA": (7, "M", None), "expTA": (8, "M", None), "sqrtPH": (9, "M", None), "invsqrtPH": (10, "M", None),
    "log": (11, "M", None), "eigenvalsH": (12, "V", None),
}
MATFUNC_LIBS = ("F3", "D3", "FN")
UNARY_OWN_FORMS = [(True, False), (False, True), (True, True)]


@dataclass(frozen=True)
class MatVariant:
    """A colour-matrix function r = f(a), or r = a^-1 b (inverse_V / inverse_M)."""
    lib: str        # "F3", "D3" or "FN" (float, leading int nc)
    fn: str         # key of MATFUNCS
    xf: bool
    pa: bool
    pb: bool = False
    ir: bool = False    # per-operand index forms (QLA_INDEX_ALL)
    ia: bool = False
    ib: bool = False

    @property
    def code(self) -> int:
        return MATFUNCS[self.fn][0]

    @property
    def dest(self) -> str:
        return MATFUNCS[self.fn][1]

    @property
    def src2(self):
        return MATFUNCS[self.fn][2]

    @property
    def own_index(self) -> bool:
        return self.ir or self.ia or self.ib

    @property
    def name(self) -> str:
        form = "" if self.own_index else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + self.dest
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + "M"
        if self.src2:
            b = ("p" if self.pb else "") + ("x" if self.ib else "") + self.src2
            return f"QLA_{self.lib}_{r}_{form}eq_{a}_inverse_{b}"
        return f"QLA_{self.lib}_{r}_{form}eq_{self.fn}_{a}"

    def params(self) -> list[tuple[str, str]]:
        fn_lib = self.lib == "FN"
        real = "D" if self.lib == "D3" else "F"

        def rec(t: str, var: str, ptr: bool) -> str:
            if t == "C":
                return f"QLA_{real}_Complex *QLA_MI355X_RESTRICT {var}"
            if fn_lib:
                return f"void *{var}"
            return f"QLA_{self.lib}_{TYPE_NAMES[t]} *QLA_MI355X_RESTRICT {'*' if ptr else ''}{var}"
        out = [("i", "int nc")] if fn_lib else []
        out.append(("p", rec(self.dest, "r", False)))
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", rec("M", "a", self.pa)))
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.src2:
            out.append(("p", rec(self.src2, "b", self.pb)))
            if self.ib:
                out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.lib == "D3" else 0
        nc = "nc" if self.lib == "FN" else "3"
        b = "(const void *)b" if self.src2 else "nullptr"
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"matfunc_entry({prec}, {self.code}, {nc}, {int(self.pa)}, {int(self.pb)}, (void *)r, "
                f"(const void *)a, {b}, {idx}, {own}, n, \"{self.name}\");")


def matfuncs() -> list[MatVariant]:
    out = []
    for lib in MATFUNC_LIBS:
        for fn in MATFUNCS:
            if MATFUNCS[fn][2]:
                for xf in (False, True):
                    for pa in (False, True):
                        for pb in (False, True):
                            out.append(MatVariant(lib, fn, xf, pa, pb))
                for ir, ia, ib in OWN_INDEX_FORMS:
                    out.append(MatVariant(lib, fn, False, False, False, ir, ia, ib))
            else:
                out += [MatVariant(lib, fn, xf, pa) for xf in (False, True) for pa in (False, True)]
                out += [MatVariant(lib, fn, False, False, False, ir, ia) for ir, ia in UNARY_OWN_FORMS]
    return out


# field linear algebra: family -> (op code of qla_launch_vec, constant ('', 'r' or 'c'),
# field sources, qualifier in the name, source transform (bit 0 conjugate, bit 1 transpose),
# record types)
SQUARE_TYPES = ("M", "P")
# the arithmetic types (datatypes.pl:64-73) without the integer fields, and
# those a complex constant multiplies (:295-302)
ARITH_TYPES = ("R", "C") + FIELD_TYPES
CMULT_TYPES = ("C",) + FIELD_TYPES
VECFAMS = {
    "copy": (0, "", 1, "", 0, ARITH_TYPES),
    "transpose": (0, "", 1, "transpose_", 2, SQUARE_TYPES),
    "conj": (0, "", 1, "conj_", 1, CMULT_TYPES),
    "i": (1, "", 1, "i_", 0, CMULT_TYPES),
    "r_times": (2, "r", 1, "", 0, ARITH_TYPES),
    "c_times": (3, "c", 1, "", 0, CMULT_TYPES),
    "plus": (4, "", 2, "", 0, ARITH_TYPES),
    "minus": (5, "", 2, "", 0, ARITH_TYPES),
    "r_times_plus": (6, "r", 2, "", 0, ARITH_TYPES),
    "r_times_minus": (7, "r", 2, "", 0, ARITH_TYPES),
    "c_times_plus": (8, "c", 2, "", 0, CMULT_TYPES),
    "c_times_minus": (9, "c", 2, "", 0, CMULT_TYPES),
}
# types with an adjoint (datatypes.pl:275-279): C (its conjugate), M, P
ADJ_TYPES = ("C",) + SQUARE_TYPES
# families with every eqop and an adjoint (`Ma`, `Pa`) source form
VEC_ALL_EQOPS = ("copy", "transpose", "conj", "i", "r_times", "c_times")
VEC_ADJ = ("copy", "r_times", "c_times")


@dataclass(frozen=True)
class VecVariant:
    """A field linear-algebra call r eqop op(x[, y]) with an optional constant."""
    prec: str       # "F3" / "D3"
    fam: str        # key of VECFAMS
    t: str          # record type letter
    eqop: int
    adj: bool = False   # copy: Hermitian conjugate; r_times / c_times: adjoint source
    xf: bool = False
    px: bool = False    # first field source is a pointer array
    py: bool = False    # second field source is a pointer array
    ir: bool = False    # per-operand index forms (QLA_INDEX_ALL)
    ix: bool = False
    iy: bool = False

    @property
    def own_index(self) -> bool:
        return self.ir or self.ix or self.iy

    @property
    def scalar(self) -> str:
        return VECFAMS[self.fam][1]

    @property
    def nsrc(self) -> int:
        return VECFAMS[self.fam][2]

    @property
    def xform(self) -> int:
        if self.adj:
            return 1 if self.t == "C" else 3
        return VECFAMS[self.fam][4]

    def _names(self) -> tuple[str, str]:
        """QLA argument names of the two field sources (prototype.pl:600-640:
        src1 a, src2 b, src3 c; a constant takes the src1 slot)."""
        return ("b", "c") if self.scalar else ("a", "b")

    @property
    def name(self) -> str:
        form = "" if self.own_index else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + self.t
        x = ("p" if self.px else "") + ("x" if self.ix else "") + self.t + ("a" if self.adj else "")
        y = ("p" if self.py else "") + ("x" if self.iy else "") + self.t
        eq = EQOPS[self.eqop]
        head = f"QLA_{lib_of(self.prec, self.t)}_{r}_{form}{eq}_"
        if self.nsrc == 1:
            if self.scalar:
                return f"{head}{self.scalar}_times_{x}"
            return f"{head}{VECFAMS[self.fam][3]}{x}"
        pm = "plus" if self.fam.endswith("plus") else "minus"
        if self.scalar:
            return f"{head}{self.scalar}_times_{x}_{pm}_{y}"
        return f"{head}{x}_{pm}_{y}"

    def params(self) -> list[tuple[str, str]]:
        rec = rec_type(self.prec, self.t)
        real = "D" if self.prec == "D3" else "F"
        xn, yn = self._names()
        out = [("p", f"{rec} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        if self.scalar:
            st = "Real" if self.scalar == "r" else "Complex"
            out.append(("p", f"QLA_{real}_{st} *QLA_MI355X_RESTRICT a"))
        out.append(("p", f"{rec} *QLA_MI355X_RESTRICT {'*' if self.px else ''}{xn}"))
        if self.ix:
            out.append(("p", f"int *index{xn}"))
        if self.nsrc == 2:
            out.append(("p", f"{rec} *QLA_MI355X_RESTRICT {'*' if self.py else ''}{yn}"))
            if self.iy:
                out.append(("p", f"int *index{yn}"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        xn, yn = self._names()
        y = f"(const void *){yn}" if self.nsrc == 2 else "nullptr"
        s = "(const void *)a" if self.scalar else "nullptr"
        sk = {"": 0, "r": 1, "c": 2}[self.scalar]
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), (f"index{xn}", self.ix),
                                                          (f"index{yn}", self.iy)))
        return (f"vec_entry({prec}, {VECFAMS[self.fam][0]}, {self.eqop}, {NCOMP[self.t]}, {self.xform}, "
                f"{int(self.px)}, {int(self.py)}, r, (const void *){xn}, {y}, {s}, {sk}, {idx}, {own}, n, "
                f"\"{self.name}\");")


def vecops() -> list[VecVariant]:
    out = []
    for prec in PRECS:
        for fam, (_, _, nsrc, _, _, types) in VECFAMS.items():
            eqops = range(4) if fam in VEC_ALL_EQOPS else (0,)
            for t in types:
                for adj in ((False, True) if fam in VEC_ADJ and t in ADJ_TYPES else (False,)):
                    for eqop in eqops:
                        if nsrc == 1:
                            out += [VecVariant(prec, fam, t, eqop, adj, xf, px) for xf in (False, True)
                                    for px in (False, True)]
                            out += [VecVariant(prec, fam, t, eqop, adj, ir=ir, ix=ix) for ir, ix in UNARY_OWN_FORMS]
                        else:
                            out += [VecVariant(prec, fam, t, eqop, adj, xf, px, py) for xf in (False, True)
                                    for px in (False, True) for py in (False, True)]
                            out += [VecVariant(prec, fam, t, eqop, adj, ir=ir, ix=ix, iy=iy)
                                    for ir, ix, iy in OWN_INDEX_FORMS]
    return out


@dataclass(frozen=True)
class FillVariant:
    """A fill: r = 0, r eqop a (one record a for every site), or for colour
    matrices r eqop c on the diagonal (build_qla.pl:2561-2620); v- and x-forms
    only (indirection.pl:163-170, @ind_fill_list)."""
    prec: str       # "F3" / "D3"
    kind: str       # "zero", "rec", "diag"
    t: str
    eqop: int
    xf: bool

    @property
    def name(self) -> str:
        form = "x" if self.xf else "v"
        eq = EQOPS[self.eqop]
        lib = lib_of(self.prec, self.t)
        if self.kind == "zero":
            return f"QLA_{lib}_{self.t}_{form}eq_zero"
        src = self.t.lower() if self.kind == "rec" else "c"
        return f"QLA_{lib}_{self.t}_{form}{eq}_{src}"

    def params(self) -> list[tuple[str, str]]:
        rec = rec_type(self.prec, self.t)
        out = [("p", f"{rec} *QLA_MI355X_RESTRICT r")]
        if self.kind == "rec":
            out.append(("p", f"{rec} *QLA_MI355X_RESTRICT a"))
        elif self.kind == "diag":
            out.append(("p", f"QLA_{'D' if self.prec == 'D3' else 'F'}_Complex *QLA_MI355X_RESTRICT a"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        kind = {"zero": 0, "rec": 1, "diag": 2}[self.kind]
        a = "nullptr" if self.kind == "zero" else "(const void *)a"
        idx = "index" if self.xf else "nullptr"
        return (f"fill_entry({prec}, {kind}, {self.eqop}, {NCOMP[self.t]}, r, {a}, {idx}, n, \"{self.name}\");")


@dataclass(frozen=True)
class GammaVariant:
    """r = Gamma_mu a (D, P) or r = a Gamma_mu (P), eq only, every unary
    indexing form (build_qla.pl:1356-1415)."""
    prec: str
    t: str
    right: bool
    xf: bool
    pa: bool
    ir: bool = False
    ia: bool = False

    @property
    def name(self) -> str:
        form = "" if (self.ir or self.ia) else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + self.t
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + self.t
        if self.right:
            return f"QLA_{self.prec}_{r}_{form}eq_{a}_times_gamma"
        return f"QLA_{self.prec}_{r}_{form}eq_gamma_times_{a}"

    def params(self) -> list[tuple[str, str]]:
        rec = f"QLA_{self.prec}_{TYPE_NAMES[self.t]}"
        out = [("p", f"{rec} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{rec} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out.append(("i", "int mu"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        own = ", ".join(x if f else "nullptr" for x, f in (("index" if self.xf else "nullptr", self.xf),
                                                          ("indexr", self.ir), ("indexa", self.ia)))
        return (f"gamma_entry({prec}, {int(self.right)}, {NCOMP[self.t]}, {int(self.pa)}, r, (const void *)a, mu, "
                f"{own}, n, \"{self.name}\");")


def gammas() -> list[GammaVariant]:
    out = []
    for prec in PRECS:
        for t, right in (("D", False), ("P", False), ("P", True)):
            out += [GammaVariant(prec, t, right, xf, pa) for xf in (False, True) for pa in (False, True)]
            out += [GammaVariant(prec, t, right, False, False, ir, ia) for ir, ia in UNARY_OWN_FORMS]
    return out


# site-local reductions and parts: kind -> (code, destination, binary, qualifier, eqops, source types)
LOCALS = {
    "norm2": (0, "R", False, "norm2_", range(4), FIELD_TYPES),
    "dot": (1, "C", True, "", range(4), FIELD_TYPES),
    "redot": (2, "R", True, "re_", range(4), FIELD_TYPES),
    "re_trace": (3, "R", False, "re_trace_", (0,), ("M",)),
    "im_trace": (4, "R", False, "im_trace_", (0,), ("M",)),
    "trace": (5, "C", False, "trace_", (0,), ("M",)),
    "antiherm": (6, "M", False, "antiherm_", (0,), ("M",)),
}


@dataclass(frozen=True)
class LocalVariant:
    """A site-local reduction (R / C record per site) or the traceless
    anti-Hermitian part of a colour matrix."""
    prec: str
    kind: str       # key of LOCALS
    t: str          # source record type
    eqop: int
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def own_index(self) -> bool:
        return self.ir or self.ia or self.ib

    @property
    def binary(self) -> bool:
        return LOCALS[self.kind][2]

    @property
    def name(self) -> str:
        _, dest, binary, qual, _, _ = LOCALS[self.kind]
        form = "" if self.own_index else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + dest
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + self.t
        head = f"QLA_{self.prec}_{r}_{form}{EQOPS[self.eqop]}_{qual}"
        if binary:
            b = ("p" if self.pb else "") + ("x" if self.ib else "") + self.t
            return f"{head}{a}_dot_{b}"
        return f"{head}{a}"

    def params(self) -> list[tuple[str, str]]:
        dest = LOCALS[self.kind][1]
        real = "D" if self.prec == "D3" else "F"
        rec = f"QLA_{self.prec}_{TYPE_NAMES[self.t]}"
        dt = {"R": f"QLA_{real}_Real", "C": f"QLA_{real}_Complex", "M": f"QLA_{self.prec}_ColorMatrix"}[dest]
        out = [("p", f"{dt} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{rec} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.binary:
            out.append(("p", f"{rec} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
            if self.ib:
                out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        b = "(const void *)b" if self.binary else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("index" if self.xf else "nullptr", self.xf),
                                                          ("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"local_entry({prec}, {LOCALS[self.kind][0]}, {self.eqop}, {NCOMP[self.t]}, {int(self.pa)}, "
                f"{int(self.pb)}, (void *)r, (const void *)a, {b}, {own}, n, \"{self.name}\");")


def locals_() -> list[LocalVariant]:
    out = []
    for prec in PRECS:
        for kind, (_, _, binary, _, eqops, types) in LOCALS.items():
            for t in types:
                for e in eqops:
                    if binary:
                        out += [LocalVariant(prec, kind, t, e, xf, pa, pb) for xf in (False, True)
                                for pa in (False, True) for pb in (False, True)]
                        out += [LocalVariant(prec, kind, t, e, ir=ir, ia=ia, ib=ib) for ir, ia, ib in OWN_INDEX_FORMS]
                    else:
                        out += [LocalVariant(prec, kind, t, e, xf, pa) for xf in (False, True) for pa in (False, True)]
                        out += [LocalVariant(prec, kind, t, e, ir=ir, ia=ia) for ir, ia in UNARY_OWN_FORMS]
    return out


# scalar-field types of the global sums (no colour): type letter -> (C type, reals per record)
SCALAR_FIELDS = {"R": ("Real", 1), "C": ("Complex", 2)}


@dataclass(frozen=True)
class SumVariant:
    """Global sum of a field, r = sum over sites of a (build_qla.pl:2449-2521):
    F3 / D3 / DF3 for V, H, D, M, P; F / D / DF for the real and complex
    fields R, C; v- and x-forms, field or pointer source."""
    lib: str        # F3 D3 DF3 (coloured) or F D DF (R, C)
    t: str
    xf: bool
    pa: bool

    @property
    def name(self) -> str:
        return f"QLA_{self.lib}_{self.t.lower()}_{'x' if self.xf else 'v'}eq_sum_{'p' if self.pa else ''}{self.t}"

    def _types(self) -> tuple[str, str]:
        out_p = "D" if self.lib in ("D3", "DF3", "D", "DF") else "F"
        in_p = "D" if self.lib in ("D3", "D") else "F"
        if self.t in SCALAR_FIELDS:
            ct = SCALAR_FIELDS[self.t][0]
            return f"QLA_{out_p}_{ct}", f"QLA_{in_p}_{ct}"
        return f"QLA_{out_p}3_{TYPE_NAMES[self.t]}", f"QLA_{in_p}3_{TYPE_NAMES[self.t]}"

    def params(self) -> list[tuple[str, str]]:
        rt, at = self._types()
        out = [("p", f"{rt} *QLA_MI355X_RESTRICT r"),
               ("p", f"{at} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a")]
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        in_d = 1 if self.lib in ("D3", "D") else 0
        out_d = 1 if self.lib in ("D3", "DF3", "D", "DF") else 0
        ncr = SCALAR_FIELDS[self.t][1] if self.t in SCALAR_FIELDS else 2 * NCOMP[self.t]
        bytes_ = ncr * (8 if out_d else 4)
        idx = "index" if self.xf else "nullptr"
        return (f"sum_entry({in_d}, {out_d}, {ncr}, {int(self.pa)}, (void *)r, (const void *)a, {idx}, n, {bytes_}, "
                f"\"{self.name}\");")


def sums() -> list[SumVariant]:
    out = []
    for lib in ("F3", "D3", "DF3"):
        for t in FIELD_TYPES:
            out += [SumVariant(lib, t, xf, pa) for xf in (False, True) for pa in (False, True)]
    for lib in ("F", "D", "DF"):
        for t in SCALAR_FIELDS:
            out += [SumVariant(lib, t, xf, pa) for xf in (False, True) for pa in (False, True)]
    return out


@dataclass(frozen=True)
class FieldVariant:
    """r eqop f * b with f a real (R) or complex (C, Ca) field and b a record
    field or its adjoint (build_qla.pl:1517-1586), or the colour outer product
    M eqop a * b^dagger of two colour vectors (:1730-1756); every binary form."""
    prec: str
    kind: str       # "R", "C" (factor field) or "outer"
    t: str          # destination / second-source record type (M for outer)
    eqop: int
    fa: bool = False    # conjugated complex factor ("Ca")
    adj: bool = False   # adjoint second source ("Ma", "Pa")
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def own_index(self) -> bool:
        return self.ir or self.ia or self.ib

    @property
    def name(self) -> str:
        form = "" if self.own_index else ("x" if self.xf else "v")
        r = ("x" if self.ir else "") + self.t
        pre_a = ("p" if self.pa else "") + ("x" if self.ia else "")
        pre_b = ("p" if self.pb else "") + ("x" if self.ib else "")
        head = f"QLA_{self.prec}_{r}_{form}{EQOPS[self.eqop]}_"
        if self.kind == "outer":
            return f"{head}{pre_a}V_times_{pre_b}Va"
        f = self.kind + ("a" if self.fa else "")
        return f"{head}{pre_a}{f}_times_{pre_b}{self.t}{'a' if self.adj else ''}"

    def params(self) -> list[tuple[str, str]]:
        real = "D" if self.prec == "D3" else "F"
        rec = f"QLA_{self.prec}_{TYPE_NAMES[self.t]}"
        if self.kind == "outer":
            at = bt = f"QLA_{self.prec}_ColorVector"
        else:
            at = f"QLA_{real}_{'Real' if self.kind == 'R' else 'Complex'}"
            bt = rec
        out = [("p", f"{rec} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{at} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out.append(("p", f"{bt} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
        if self.ib:
            out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        if self.kind == "outer":
            return (f"outer_entry({prec}, {self.eqop}, {int(self.pa)}, {int(self.pb)}, (void *)r, (const void *)a, "
                    f"(const void *)b, {idx}, {own}, n, \"{self.name}\");")
        return (f"field_entry({prec}, {int(self.kind == 'C')}, {self.eqop}, {NCOMP[self.t]}, {3 if self.adj else 0}, "
                f"{int(self.fa)}, {int(self.pa)}, {int(self.pb)}, (void *)r, (const void *)a, (const void *)b, {idx}, "
                f"{own}, n, \"{self.name}\");")


def fieldmuls() -> list[FieldVariant]:
    out = []

    def forms(**kw):
        res = [FieldVariant(xf=xf, pa=pa, pb=pb, **kw) for xf in (False, True) for pa in (False, True)
               for pb in (False, True)]
        res += [FieldVariant(ir=ir, ia=ia, ib=ib, **kw) for ir, ia, ib in OWN_INDEX_FORMS]
        return res
    for prec in PRECS:
        for kind in ("R", "C"):
            for fa in ((False,) if kind == "R" else (False, True)):
                for t in FIELD_TYPES:
                    for adj in ((False, True) if t in SQUARE_TYPES else (False,)):
                        for e in range(4):
                            out += forms(prec=prec, kind=kind, t=t, eqop=e, fa=fa, adj=adj)
        for e in range(4):
            out += forms(prec=prec, kind="outer", t="M", eqop=e)
    return out


# propagator products: kind -> (first, second source record types)
PMUL_KINDS = {0: ("M", "P"), 1: ("P", "M"), 2: ("P", "P")}


@dataclass(frozen=True)
class PmulVariant:
    """P eqop a * b with a colour matrix times a propagator (build_qla.pl:
    1587-1620), a propagator times a colour matrix or two propagators
    (:1622-1662); either factor may be its adjoint; every binary form."""
    prec: str
    kind: int
    eqop: int
    ca: bool = False
    cb: bool = False
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def own_index(self) -> bool:
        return self.ir or self.ia or self.ib

    @property
    def name(self) -> str:
        form = "" if self.own_index else ("x" if self.xf else "v")
        ta, tb = PMUL_KINDS[self.kind]
        pre_a = ("p" if self.pa else "") + ("x" if self.ia else "")
        pre_b = ("p" if self.pb else "") + ("x" if self.ib else "")
        return (f"QLA_{self.prec}_{'x' if self.ir else ''}P_{form}{EQOPS[self.eqop]}_{pre_a}{ta}{'a' if self.ca else ''}"
                f"_times_{pre_b}{tb}{'a' if self.cb else ''}")

    def params(self) -> list[tuple[str, str]]:
        ta, tb = PMUL_KINDS[self.kind]
        out = [("p", f"QLA_{self.prec}_DiracPropagator *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"QLA_{self.prec}_{TYPE_NAMES[ta]} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out.append(("p", f"QLA_{self.prec}_{TYPE_NAMES[tb]} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
        if self.ib:
            out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"pmul_entry({prec}, {self.kind}, {self.eqop}, {int(self.ca)}, {int(self.cb)}, {int(self.pa)}, "
                f"{int(self.pb)}, (void *)r, (const void *)a, (const void *)b, {idx}, {own}, n, \"{self.name}\");")


def pmuls() -> list[PmulVariant]:
    out = []
    for prec in PRECS:
        for kind in PMUL_KINDS:
            for ca in (False, True):
                for cb in (False, True):
                    for e in range(4):
                        kw = dict(prec=prec, kind=kind, eqop=e, ca=ca, cb=cb)
                        out += [PmulVariant(xf=xf, pa=pa, pb=pb, **kw) for xf in (False, True)
                                for pa in (False, True) for pb in (False, True)]
                        out += [PmulVariant(ir=ir, ia=ia, ib=ib, **kw) for ir, ia, ib in OWN_INDEX_FORMS]
    return out


# component access: qualifier -> (moved part, array types, kind code of the get form)
COMP_FAMS = {"elem": ("C", ("M", "H", "D", "V", "P"), 0), "colorvec": ("V", ("M", "H", "D", "P"), 2),
             "diracvec": ("D", ("P",), 4), "spintrace": ("M", ("P",), 6)}
COMP_TCODE = {"V": 0, "H": 1, "D": 2, "M": 3, "P": 4}
# the reference's extra arguments (build_qla.pl:786-790, 841-845, 897-899)
ELEM_ARGS = {"V": ("ic",), "H": ("ic", "is"), "D": ("ic", "is"), "M": ("ic", "jc"), "P": ("ic", "is", "jc", "js")}
CVEC_ARGS = {"M": ("jc",), "H": ("is",), "D": ("is",), "P": ("is", "jc", "js")}


@dataclass(frozen=True)
class CompVariant:
    """Component access, eq only, every unary indexing form: an element
    (C <-> T; build_qla.pl:767-820), a colour column (V <-> M, H, D, P;
    :822-876), a Dirac column (D <-> P; :878-932), and the spin trace
    M = spintrace(P) (:1066-1086)."""
    prec: str
    qual: str
    t: str          # the array type
    get: bool
    xf: bool = False
    pa: bool = False
    ir: bool = False
    ia: bool = False

    @property
    def extra(self) -> tuple[str, ...]:
        if self.qual == "elem":
            return ELEM_ARGS[self.t]
        if self.qual == "colorvec":
            return CVEC_ARGS[self.t]
        if self.qual == "diracvec":
            return ("jc", "js")
        return ()

    @property
    def types(self) -> tuple[str, str]:
        part = COMP_FAMS[self.qual][0]
        return (part, self.t) if self.get else (self.t, part)

    @property
    def name(self) -> str:
        form = "" if (self.ir or self.ia) else ("x" if self.xf else "v")
        d, s = self.types
        return (f"QLA_{self.prec}_{'x' if self.ir else ''}{d}_{form}eq_{self.qual}_{'p' if self.pa else ''}"
                f"{'x' if self.ia else ''}{s}")

    def _ctype(self, x: str) -> str:
        if x == "C":
            return f"QLA_{'D' if self.prec == 'D3' else 'F'}_Complex"
        return f"QLA_{self.prec}_{TYPE_NAMES[x]}"

    def params(self) -> list[tuple[str, str]]:
        d, s = self.types
        out = [("p", f"{self._ctype(d)} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{self._ctype(s)} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out += [("i", f"int {x}") for x in self.extra]
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        kind = COMP_FAMS[self.qual][2] + (0 if self.get else 1)
        e = list(self.extra) + ["0"] * (4 - len(self.extra))
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia)))
        return (f"comp_entry({prec}, {kind}, {COMP_TCODE[self.t]}, {', '.join(e)}, {int(self.pa)}, (void *)r, "
                f"(const void *)a, {idx}, {own}, n, \"{self.name}\");")


def comps() -> list[CompVariant]:
    out = []
    for prec in PRECS:
        for qual, (_, types, _) in COMP_FAMS.items():
            for t in types:
                for get in ((True,) if qual == "spintrace" else (True, False)):
                    kw = dict(prec=prec, qual=qual, t=t, get=get)
                    out += [CompVariant(xf=xf, pa=pa, **kw) for xf in (False, True) for pa in (False, True)]
                    out += [CompVariant(ir=ir, ia=ia, **kw) for ir, ia in UNARY_OWN_FORMS]
    return out


# real / complex conversions and division: kind -> (destination, sources, name form)
SCONV = {0: ("C", ("R",), "{a}"), 1: ("C", ("R", "R"), "{a}_plus_i_{b}"), 2: ("R", ("C",), "re_{a}"),
         3: ("R", ("C",), "im_{a}"), 4: ("R", ("R", "R"), "{a}_divide_{b}"), 5: ("C", ("C", "C"), "{a}_divide_{b}"),
         6: ("R", ("I",), "{a}"), 7: ("I", ("R",), "trunc_{a}"), 8: ("I", ("R",), "round_{a}")}


@dataclass(frozen=True)
class SconvVariant:
    """Conversions between real and complex fields (build_qla.pl:623-708) and
    the division of real and complex fields (:1457-1482), eq only, every
    indexing form."""
    prec: str
    kind: int
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def binary(self) -> bool:
        return len(SCONV[self.kind][1]) == 2

    @property
    def name(self) -> str:
        dest, srcs, fmt = SCONV[self.kind]
        form = "" if (self.ir or self.ia or self.ib) else ("x" if self.xf else "v")
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + srcs[0]
        b = (("p" if self.pb else "") + ("x" if self.ib else "") + srcs[1]) if self.binary else ""
        return f"QLA_{self.prec[0]}_{'x' if self.ir else ''}{dest}_{form}eq_{fmt.format(a=a, b=b)}"

    def params(self) -> list[tuple[str, str]]:
        dest, srcs, _ = SCONV[self.kind]
        out = [("p", f"{rec_type(self.prec, dest)} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{rec_type(self.prec, srcs[0])} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.binary:
            out.append(("p", f"{rec_type(self.prec, srcs[1])} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
            if self.ib:
                out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        b = "(const void *)b" if self.binary else "nullptr"
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"sconv_entry({prec}, {self.kind}, {int(self.pa)}, {int(self.pb)}, (void *)r, (const void *)a, {b}, "
                f"{idx}, {own}, n, \"{self.name}\");")


def sconvs() -> list[SconvVariant]:
    out = []
    for prec in PRECS:
        for kind in SCONV:
            kw = dict(prec=prec, kind=kind)
            if len(SCONV[kind][1]) == 2:
                out += [SconvVariant(xf=xf, pa=pa, pb=pb, **kw) for xf in (False, True) for pa in (False, True)
                        for pb in (False, True)]
                out += [SconvVariant(ir=ir, ia=ia, ib=ib, **kw) for ir, ia, ib in OWN_INDEX_FORMS]
            else:
                out += [SconvVariant(xf=xf, pa=pa, **kw) for xf in (False, True) for pa in (False, True)]
                out += [SconvVariant(ir=ir, ia=ia, **kw) for ir, ia in UNARY_OWN_FORMS]
    return out


# elementary functions of real and complex fields: name -> (code, destination, sources, name form)
SFN_UNARY = ("cos", "sin", "tan", "acos", "asin", "atan", "sqrt", "fabs", "exp", "log", "sign", "cosh", "sinh",
             "tanh", "log10", "floor", "ceil")
SFN = {f: (k, "R", ("R",), f + "_{a}") for k, f in enumerate(SFN_UNARY)}
SFN.update({op: (20 + k, "R", ("R", "I" if op == "ldexp" else "R"), "{a}_" + op + "_{b}")
            for k, op in enumerate(("mod", "max", "min", "pow", "atan2", "ldexp"))})
SFN.update({"cexpi": (30, "C", ("R",), "cexpi_{a}"), "norm": (31, "R", ("C",), "norm_{a}"),
            "arg": (32, "R", ("C",), "arg_{a}"), "cexp": (33, "C", ("C",), "cexp_{a}"),
            "csqrt": (34, "C", ("C",), "csqrt_{a}"), "clog": (35, "C", ("C",), "clog_{a}")})


@dataclass(frozen=True)
class SfnVariant:
    """Elementary functions of real and complex fields, eq only, every
    indexing form (build_qla.pl:194-355)."""
    prec: str
    fn: str
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def binary(self) -> bool:
        return len(SFN[self.fn][2]) == 2

    @property
    def name(self) -> str:
        _, dest, srcs, fmt = SFN[self.fn]
        form = "" if (self.ir or self.ia or self.ib) else ("x" if self.xf else "v")
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + srcs[0]
        b = (("p" if self.pb else "") + ("x" if self.ib else "") + srcs[1]) if self.binary else ""
        return f"QLA_{self.prec[0]}_{'x' if self.ir else ''}{dest}_{form}eq_{fmt.format(a=a, b=b)}"

    def params(self) -> list[tuple[str, str]]:
        _, dest, srcs, _ = SFN[self.fn]
        ct = lambda x: "QLA_Int" if x == "I" else rec_type(self.prec, x)   # noqa: E731
        out = [("p", f"{ct(dest)} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{ct(srcs[0])} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        if self.binary:
            out.append(("p", f"{ct(srcs[1])} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
            if self.ib:
                out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        b = "(const void *)b" if self.binary else "nullptr"
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"sfn_entry({prec}, {SFN[self.fn][0]}, {int(self.pa)}, {int(self.pb)}, (void *)r, (const void *)a, "
                f"{b}, {idx}, {own}, n, \"{self.name}\");")


def sfns() -> list[SfnVariant]:
    out = []
    for prec in PRECS:
        for fn, (_, _, srcs, _) in SFN.items():
            kw = dict(prec=prec, fn=fn)
            if len(srcs) == 2:
                out += [SfnVariant(xf=xf, pa=pa, pb=pb, **kw) for xf in (False, True) for pa in (False, True)
                        for pb in (False, True)]
                out += [SfnVariant(ir=ir, ia=ia, ib=ib, **kw) for ir, ia, ib in OWN_INDEX_FORMS]
            else:
                out += [SfnVariant(xf=xf, pa=pa, **kw) for xf in (False, True) for pa in (False, True)]
                out += [SfnVariant(ir=ir, ia=ia, **kw) for ir, ia in UNARY_OWN_FORMS]
    return out


CMP_OPS = ("eq", "ne", "gt", "lt", "ge", "le")


@dataclass(frozen=True)
class MaskVariant:
    """Copymask r = a where the int field b is nonzero, for every floating-point
    type (build_qla.pl:2058-2088), and the comparisons of real fields
    I = R op R (:1976-2018); eq only, every binary form."""
    prec: str
    t: str              # copied type, or "cmp"
    op: str = ""        # comparison
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def name(self) -> str:
        form = "" if (self.ir or self.ia or self.ib) else ("x" if self.xf else "v")
        pre = lambda ptr, own: ("p" if ptr else "") + ("x" if own else "")   # noqa: E731
        if self.t == "cmp":
            return (f"QLA_{self.prec[0]}_{'x' if self.ir else ''}I_{form}eq_{pre(self.pa, self.ia)}R_{self.op}_"
                    f"{pre(self.pb, self.ib)}R")
        return (f"QLA_{lib_of(self.prec, self.t)}_{'x' if self.ir else ''}{self.t}_{form}eq_"
                f"{pre(self.pa, self.ia)}{self.t}_mask_{pre(self.pb, self.ib)}I")

    def params(self) -> list[tuple[str, str]]:
        if self.t == "cmp":
            dt, at, bt = "QLA_Int", rec_type(self.prec, "R"), rec_type(self.prec, "R")
        else:
            dt = at = rec_type(self.prec, self.t)
            bt = "QLA_Int"
        out = [("p", f"{dt} *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        out.append(("p", f"{at} *QLA_MI355X_RESTRICT {'*' if self.pa else ''}a"))
        if self.ia:
            out.append(("p", "int *indexa"))
        out.append(("p", f"{bt} *QLA_MI355X_RESTRICT {'*' if self.pb else ''}b"))
        if self.ib:
            out.append(("p", "int *indexb"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        prec = 1 if self.prec == "D3" else 0
        op = CMP_OPS.index(self.op) if self.t == "cmp" else -1
        nw = 0 if self.t == "cmp" else (1 if self.t == "R" else 2 * NCOMP[self.t])
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", self.ir), ("indexa", self.ia),
                                                          ("indexb", self.ib)))
        return (f"mask_entry({prec}, {op}, {nw}, {int(self.pa)}, {int(self.pb)}, (void *)r, (const void *)a, "
                f"(const void *)b, {idx}, {own}, n, \"{self.name}\");")


def masks() -> list[MaskVariant]:
    out = []
    for prec in PRECS:
        for t, op in [(t, "") for t in ARITH_TYPES] + [("cmp", o) for o in CMP_OPS]:
            kw = dict(prec=prec, t=t, op=op)
            out += [MaskVariant(xf=xf, pa=pa, pb=pb, **kw) for xf in (False, True) for pa in (False, True)
                    for pb in (False, True)]
            out += [MaskVariant(ir=ir, ia=ia, ib=ib, **kw) for ir, ia, ib in OWN_INDEX_FORMS]
    return out


# integer fields: name -> (kernel op, binary, every eqop, int constant, name form)
INT_OPS = {"copy": (0, False, True, False, "{a}"), "i_times": (1, False, True, True, "i_times_{a}"),
           "not": (2, False, False, False, "not_{a}"), "times": (13, True, True, False, "{a}_times_{b}"),
           "i_times_plus": (14, True, False, True, "i_times_{a}_plus_{b}"),
           "i_times_minus": (15, True, False, True, "i_times_{a}_minus_{b}")}
INT_OPS.update({op: (code, True, False, False, "{a}_" + op + "_{b}") for code, op in (
    (10, "plus"), (11, "minus"), (12, "divide"), (16, "lshift"), (17, "rshift"), (18, "mod"), (19, "max"),
    (20, "min"), (21, "eq"), (22, "ne"), (23, "gt"), (24, "lt"), (25, "ge"), (26, "le"), (27, "or"), (28, "and"),
    (29, "xor"), (30, "mask"))})


@dataclass(frozen=True)
class IntVariant:
    """The integer fields (no precision or colour: QLA_ prefix only,
    prototype.pl:327-353): copies, the int constant, sums, division,
    multiplication, shifts, mod, max, min, comparisons, boolean operations,
    copymask (build_qla.pl:356-402, 1256-1294, 1421-1516, 1837-1888,
    1936-2088) and the fills (:2561-2620); every form."""
    op: str             # key of INT_OPS, or "zero" / "fill"
    eqop: int = 0
    xf: bool = False
    pa: bool = False
    pb: bool = False
    ir: bool = False
    ia: bool = False
    ib: bool = False

    @property
    def fill(self) -> bool:
        return self.op in ("zero", "fill")

    @property
    def binary(self) -> bool:
        return not self.fill and INT_OPS[self.op][1]

    @property
    def scalar(self) -> bool:
        return self.op == "fill" or (not self.fill and INT_OPS[self.op][3])

    @property
    def name(self) -> str:
        form = "" if (self.ir or self.ia or self.ib) else ("x" if self.xf else "v")
        head = f"QLA_{'x' if self.ir else ''}I_{form}{EQOPS[self.eqop]}_"
        if self.op == "zero":
            return head + "zero"
        if self.op == "fill":
            return head + "i"
        a = ("p" if self.pa else "") + ("x" if self.ia else "") + "I"
        b = ("p" if self.pb else "") + ("x" if self.ib else "") + "I"
        return head + INT_OPS[self.op][4].format(a=a, b=b)

    def _names(self) -> tuple[str, str]:
        return ("b", "c") if self.scalar else ("a", "b")

    def params(self) -> list[tuple[str, str]]:
        out = [("p", "QLA_Int *QLA_MI355X_RESTRICT r")]
        if self.ir:
            out.append(("p", "int *indexr"))
        if self.scalar:
            out.append(("p", "QLA_Int *QLA_MI355X_RESTRICT a"))
        if not self.fill:
            xn, yn = self._names()
            out.append(("p", f"QLA_Int *QLA_MI355X_RESTRICT {'*' if self.pa else ''}{xn}"))
            if self.ia:
                out.append(("p", f"int *index{xn}"))
            if self.binary:
                out.append(("p", f"QLA_Int *QLA_MI355X_RESTRICT {'*' if self.pb else ''}{yn}"))
                if self.ib:
                    out.append(("p", f"int *index{yn}"))
        if self.xf:
            out.append(("p", "int *index"))
        out.append(("i", "int n"))
        return out

    def ctypes_kinds(self) -> str:
        return "".join(k for k, _ in self.params())

    def body(self) -> str:
        code = 40 if self.op == "zero" else 41 if self.op == "fill" else INT_OPS[self.op][0]
        xn, yn = self._names()
        s = "(const void *)a" if self.scalar else "nullptr"
        x = "nullptr" if self.fill else f"(const void *){xn}"
        y = f"(const void *){yn}" if self.binary else "nullptr"
        idx = "index" if self.xf else "nullptr"
        own = ", ".join(v if f else "nullptr" for v, f in (("indexr", self.ir), (f"index{xn}", self.ia),
                                                          (f"index{yn}", self.ib)))
        return (f"int_entry({code}, {self.eqop}, {int(self.pa)}, {int(self.pb)}, (void *)r, {s}, {x}, {y}, {idx}, "
                f"{own}, n, \"{self.name}\");")


def ints() -> list[IntVariant]:
    out = [IntVariant("zero", 0, xf) for xf in (False, True)]
    out += [IntVariant("fill", e, xf) for e in range(4) for xf in (False, True)]
    for op, (_, binary, alleq, _, _) in INT_OPS.items():
        for e in (range(4) if alleq else (0,)):
            if binary:
                out += [IntVariant(op, e, xf, pa, pb) for xf in (False, True) for pa in (False, True)
                        for pb in (False, True)]
                out += [IntVariant(op, e, ir=ir, ia=ia, ib=ib) for ir, ia, ib in OWN_INDEX_FORMS]
            else:
                out += [IntVariant(op, e, xf, pa) for xf in (False, True) for pa in (False, True)]
                out += [IntVariant(op, e, ir=ir, ia=ia) for ir, ia in UNARY_OWN_FORMS]
    return out


def fills() -> list[FillVariant]:
    out = []
    for prec in PRECS:
        for t in ARITH_TYPES:
            out += [FillVariant(prec, "zero", t, 0, xf) for xf in (False, True)]
        for t in ARITH_TYPES:
            out += [FillVariant(prec, "rec", t, e, xf) for e in range(4) for xf in (False, True)]
        out += [FillVariant(prec, "diag", "M", e, xf) for e in range(4) for xf in (False, True)]
    return out


def reductions() -> list[RedVariant]:
    out = []
    for prec in ("F3", "D3", "DF3"):
        for kind in ("norm2", "dot", "redot"):
            for t in FIELD_TYPES:
                for xf in (False, True):
                    for pa in (False, True):
                        for pb in ((False,) if kind == "norm2" else (False, True)):
                            out.append(RedVariant(prec, kind, t, xf, pa, pb))
                if kind != "norm2":
                    for ia, ib in ((True, False), (False, True), (True, True)):
                        out.append(RedVariant(prec, kind, t, False, False, False, ia, ib))
    return out


def conversions() -> list[ConvVariant]:
    out = []
    for lib in ("DF3", "FD3"):
        for t in FIELD_TYPES:
            out += [ConvVariant(lib, t, xf, pa) for xf in (False, True) for pa in (False, True)]
            out += [ConvVariant(lib, t, False, False, ir, ia) for ir, ia in ((True, False), (False, True), (True, True))]
    return out


def all_symbols() -> list:
    """Every generated symbol (site family, reductions, conversions, matrix functions,
    field linear algebra, fills, gamma multiplications, site-local reductions, global sums, multiplication
    by real and complex fields, outer products, products with propagators, component access, spin trace)."""
    return (variants() + reductions() + conversions() + matfuncs() + vecops() + fills() + gammas() + locals_()
            + sums() + fieldmuls() + pmuls() + comps() + sconvs() + sfns() + masks() + ints())


def variants() -> list[Variant]:
    out = []
    for prec in PRECS:
        for kind in KINDS:
            if kind in UNARY_KINDS:
                for xf in (False, True):
                    for eqop in range(4):
                        for pa in (False, True):
                            out.append(Variant(prec, kind, eqop, False, xf, pa, False))
                for ir, ia in UNARY_OWN_FORMS:
                    for eqop in range(4):
                        out.append(Variant(prec, kind, eqop, False, False, False, False, ir, ia))
                continue
            for xf in (False, True):
                for eqop in range(4):
                    for adj in (False, True):
                        for pa in (False, True):
                            for pb in (False, True):
                                out.append(Variant(prec, kind, eqop, adj, xf, pa, pb))
            # per-operand index forms (QLA_INDEX_ALL): fields only, no gang index
            for ir, ia, ib in OWN_INDEX_FORMS:
                for eqop in range(4):
                    for adj in (False, True):
                        out.append(Variant(prec, kind, eqop, adj, False, False, False, ir, ia, ib))
    return out


def _decl(v: Variant) -> str:
    ps = [d for _, d in v.params()]
    head = f"void {v.name}("
    pad = " " * len(head)
    return head + (",\n" + pad).join(ps) + ")"


_FAMILY_DOC = {
    "matvec": ("colour matrix times colour vector, r[ic] eqop sum_kc A[ic][kc] b[kc]",
               "lib/perl/expressions_tensor.pl:544-630 (x=0 or x=r, then += / -= over kc)"),
    "spproj": ("spin projection of B, then colour matrix times the HalfFermion",
               "build_qla.pl:1162-1190, make_code_binary.pl:189-222,259-271, expressions_gamma.pl:244-276"),
    "sprecon": ("colour matrix times B, then spin reconstruction into a DiracFermion",
                "build_qla.pl:1192-1220, make_code_binary.pl:224-257, expressions_gamma.pl:311-337"),
    "wilson": ("spin projection of B, colour matrix times it, spin reconstruction",
               "build_qla.pl:1222-1250, make_code_binary.pl:285-341"),
    "multi": ("r eqop sum_{d<nd} A_d * B_d (eqop for d = 0, then += / -=; nd <= 0 is a no-op)",
              "build_qla.pl:1693-1725, expressions_tensor.pl:983-1002, prototype.pl:683-685"),
    "mat_h": ("colour matrix times HalfFermion, r[ic][is] eqop sum_kc A[ic][kc] b[kc][is]",
              "build_qla.pl:1587-1622, expressions_tensor.pl:544-630"),
    "mat_d": ("colour matrix times DiracFermion, r[ic][is] eqop sum_kc A[ic][kc] b[kc][is]",
              "build_qla.pl:1587-1622, expressions_tensor.pl:544-630"),
    "mat_m": ("colour matrix times colour matrix, r[ic][jc] eqop sum_kc A[ic][kc] b[kc][jc]",
              "build_qla.pl:1587-1622, expressions_tensor.pl:544-630"),
    "mat_ma": ("colour matrix times adjoint colour matrix, r[ic][jc] eqop sum_kc A[ic][kc] conj(b[jc][kc]) "
               "(QLA_c_peq_c_times_ca / _ca_times_ca)", "build_qla.pl:1587-1622, expressions_tensor.pl:544-630"),
    "proj": ("spin projection of a DiracFermion, r eqop spproj(a)",
             "build_qla.pl:1087-1110, make_code_unary.pl:183-202, expressions_gamma.pl:205-276"),
    "recon": ("spin reconstruction of a HalfFermion, r eqop sprecon(a)",
              "build_qla.pl:1112-1135, make_code_unary.pl:204-223, expressions_gamma.pl:278-337"),
    "projrecon": ("spin projection with reconstruction, r eqop sprecon(spproj(a))",
                  "build_qla.pl:1137-1160, make_code_unary.pl:243-283"),
}


def header_text() -> str:
    vs = variants()
    lines = [
        "/*",
        " * qla_mi355x_sites.h -- Wilson-hop building blocks of the QLA C ABI on MI355X.",
        " * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        " * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device",
        " * semantics are documented there).",
        " *",
        " * Every symbol keeps the name and argument order of the QLA 1.9.0 symbol",
        " * (lib/perl/prototype.pl:450-688): dest, src1, src2, `int mu, int sign`",
        " * for the spin forms, the gang `int *index` for x-forms, `int n`, and",
        " * `int nd` for the multi-source forms.  `Ma` uses the adjoint of the",
        " * matrix; a `p` operand is a pointer array (device array of device record",
        " * addresses); x-forms use every operand at index[i].  Multi-source forms:",
        " * `a` and `b` are arrays of nd field (or pointer-array) addresses held in",
        " * host or device memory.",
        " */",
        "#ifndef QLA_MI355X_SITES_H",
        "#define QLA_MI355X_SITES_H",
    ]
    last = None
    for v in vs:
        key = (v.kind, v.prec)
        if key != last:
            what, ref = _FAMILY_DOC[v.kind]
            lines += ["", f"/* {v.prec}: {what}.", f" * Reference: {ref}. */"]
            last = key
        lines.append(_decl(v) + ";")
    lines += ["", "#endif", ""]
    return "\n".join(lines)


def inc_text() -> str:
    lines = [
        "// qla_sites_abi.inc -- definitions of the Wilson-hop building blocks of the C ABI.",
        "// GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        "// Included by qla_capi.hip inside extern \"C\"; site_entry / multi_entry are there.",
    ]
    for v in variants():
        prec = 0 if v.prec == "F3" else 1
        lines.append(_decl(v))
        idx = "index" if v.xf else "nullptr"
        own = ", ".join(x if f else "nullptr" for x, f in (("indexr", v.ir), ("indexa", v.ia), ("indexb", v.ib)))
        if v.unary:
            own_u = ", ".join(x if f else "nullptr" for x, f in (("indexr", v.ir), ("nullptr", False), ("indexa", v.ia)))
            body = (f"site_entry({prec}, {KINDS[v.kind]}, {v.eqop}, 0, 0, {int(v.pa)}, r, nullptr, "
                    f"(const void *)a, mu, sign, {idx}, {own_u}, n, \"{v.name}\");")
        elif v.multi:
            body = (f"multi_entry({prec}, {v.eqop}, {int(v.adj)}, {int(v.pa)}, {int(v.pb)}, r, "
                    f"(const void *)a, (const void *)b, {idx}, {own}, n, nd, \"{v.name}\");")
        else:
            mu, sign = ("mu", "sign") if v.spin else ("0", "1")
            body = (f"site_entry({prec}, {KINDS[v.kind]}, {v.eqop}, {int(v.adj)}, {int(v.pa)}, {int(v.pb)}, r, "
                    f"(const void *)a, (const void *)b, {mu}, {sign}, {idx}, {own}, n, \"{v.name}\");")
        lines.append("{")
        lines.append("    " + body)
        lines.append("}")
    lines.append("")
    return "\n".join(lines)


_FIELDS_DOC = {
    "norm2": ("global squared norm, sum over sites and components of |a|^2",
              "build_qla.pl:2089-2156, make_code_unary.pl:318-366; per component the norm is formed "
              "in the field precision and added to a double sum"),
    "dot": ("global inner product, sum of conj(a) b over sites and components",
            "build_qla.pl:2196-2360, make_code_binary.pl:110-170; operands promoted to double"),
    "redot": ("real part of the global inner product",
              "build_qla.pl:2244-2260 (qualifier re), make_code_binary.pl:110-170"),
    "conv": ("precision conversion of a field (DF3: float to double, exact; FD3: double to float, "
             "rounded to nearest)", "build_qla.pl:530-585, make_code_unary.pl (identity with a cast)"),
}


def fields_header_text() -> str:
    lines = [
        "/*",
        " * qla_mi355x_fields.h -- global reductions and precision conversions of the QLA",
        " * C ABI on MI355X (types V, H, D, M, P).",
        " * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        " * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device",
        " * semantics are documented there).",
        " *",
        " * Names and argument order are QLA's (lib/perl/prototype.pl:450-688).  A `p`",
        " * operand is a pointer array; x-forms use the records at index[i].  The result",
        " * pointer of a reduction may be a host or a device address; n <= 0 gives 0.",
        " * QLA_DF3_* reductions take float fields and return double results; QLA_DF3_*",
        " * / QLA_FD3_* conversions write double / float fields.",
        " */",
        "#ifndef QLA_MI355X_FIELDS_H",
        "#define QLA_MI355X_FIELDS_H",
    ]
    last = None
    for v in reductions() + conversions():
        kind = v.kind if isinstance(v, RedVariant) else "conv"
        lib = v.prec if isinstance(v, RedVariant) else v.lib
        key = (kind, lib)
        if key != last:
            what, ref = _FIELDS_DOC[kind]
            lines += ["", f"/* {lib}: {what}.", f" * Reference: {ref}. */"]
            last = key
        lines.append(_decl(v) + ";")
    lines += ["", "#endif", ""]
    return "\n".join(lines)


def fields_inc_text() -> str:
    lines = [
        "// qla_fields_abi.inc -- definitions of the global reductions and precision conversions.",
        "// GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        "// Included by qla_capi.hip inside extern \"C\"; reduce_entry / convert_entry are there.",
    ]
    for v in reductions() + conversions():
        lines += [_decl(v), "{", "    " + v.body(), "}"]
    lines.append("")
    return "\n".join(lines)


_MATFUNC_DOC = {
    "det": ("determinant", "code-templates/C_eq_det_M.c (closed form for nc = 3, pivoted LU otherwise)"),
    "inverse": ("matrix inverse", "code-templates/M_eq_inverse_M.c (adjugate for nc = 3, Gauss-Jordan otherwise)"),
    "inverse_V": ("a^-1 b for a colour vector b",
                  "code-templates/V_eq_M_inverse_V.c (adjugate for nc = 3, Gauss-Jordan otherwise)"),
    "inverse_M": ("a^-1 b for a colour matrix b", "code-templates/M_eq_M_inverse_M.c (Gauss-Jordan)"),
    "exp": ("matrix exponential", "code-templates/M_eq_exp_M.c (scaled Pade(3,3), squaring)"),
    "expA": ("exponential of an anti-Hermitian matrix",
             "code-templates/M_eq_expA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise)"),
    "expTA": ("exponential of a traceless anti-Hermitian matrix",
              "code-templates/M_eq_expTA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise)"),
    "sqrt": ("matrix square root", "code-templates/M_eq_sqrt_M.c (Newton iteration)"),
    "sqrtPH": ("square root of a positive Hermitian matrix",
               "code-templates/M_eq_sqrtPH_M.c (eigenvalue closed form for nc = 3, the sqrt branch otherwise)"),
    "invsqrt": ("inverse square root", "code-templates/M_eq_invsqrt_M.c (Newton iteration)"),
    "invsqrtPH": ("inverse square root of a positive Hermitian matrix",
                  "code-templates/M_eq_invsqrtPH_M.c (closed form for nc = 3, the invsqrt branch otherwise)"),
    "log": ("matrix logarithm", "code-templates/M_eq_log_M.c (4 square roots, Pade, rescale)"),
    "eigenvalsH": ("eigenvalues of a Hermitian matrix (nc = 3: closed form, template order, imaginary parts 0; "
                   "QLA_FN nc = 4..8: the template's QR iteration)",
                   "code-templates/V_eq_eigenvalsH_M.c:103-161 (nc = 3), :180-484 (QR)"),
}


def matfuncs_header_text() -> str:
    lines = [
        "/*",
        " * qla_mi355x_matfuncs.h -- colour-matrix functions of the QLA C ABI on MI355X.",
        " * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        " * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device",
        " * semantics are documented there).",
        " *",
        " * Names and argument order are QLA's (lib/perl/build_qla.pl:1002-1066 and",
        " * 1663-1690, prototype.pl:450-688).  QLA_FN_* take the number of colours",
        " * first (nc = 3..8 here; other values record an error and write nothing)",
        " * and their record arguments are untyped: nc x nc complex floats per matrix,",
        " * nc per vector, or a pointer array of such records for a `p` operand.  A `p`",
        " * operand is a pointer array; x-forms use every operand at index[i].",
        " */",
        "#ifndef QLA_MI355X_MATFUNCS_H",
        "#define QLA_MI355X_MATFUNCS_H",
    ]
    last = None
    for v in matfuncs():
        key = (v.lib, v.fn)
        if key != last:
            what, ref = _MATFUNC_DOC[v.fn]
            lines += ["", f"/* {v.lib}: {what}.", f" * Reference: {ref}. */"]
            last = key
        lines.append(_decl(v) + ";")
    lines += ["", "#endif", ""]
    return "\n".join(lines)


def matfuncs_inc_text() -> str:
    lines = [
        "// qla_matfuncs_abi.inc -- definitions of the colour-matrix functions of the C ABI.",
        "// GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        "// Included by qla_capi.hip inside extern \"C\"; matfunc_entry is there.",
    ]
    for v in matfuncs():
        lines += [_decl(v), "{", "    " + v.body(), "}"]
    lines.append("")
    return "\n".join(lines)


_VEC_DOC = {
    "copy": ("copy or increment, r eqop a; `Ta` is the Hermitian conjugate",
             "build_qla.pl:356-430, macros QLA_c_<eqop>_c / _ca (include/qla_complex.h:147-163)"),
    "transpose": ("transpose, r eqop a^T (P: the (colour, spin) pairs swapped)",
                  "build_qla.pl:432-460, lib/perl/headers.pl:68-90"),
    "conj": ("complex conjugate, r eqop conj(a)", "build_qla.pl:462-493, QLA_c_<eqop>_ca"),
    "i": ("times i, r eqop i a", "build_qla.pl:1328-1354, QLA_c_<eqop>_ic (include/qla_complex.h:165-172)"),
    "r_times": ("times a real constant, r eqop a[0] b (`Ta`: the adjoint of b)",
                "build_qla.pl:1256-1293, QLA_c_<eqop>_r_times_c / _ca (include/qla_complex.h:259-284)"),
    "c_times": ("times a complex constant, r eqop a[0] b (`Ta`: the adjoint of b)",
                "build_qla.pl:1295-1326, QLA_c_<eqop>_c_times_c / _ca (include/qla_complex.h:326-379)"),
    "plus": ("sum, r = a + b", "build_qla.pl:1421-1456, QLA_c_eq_c_plus_c"),
    "minus": ("difference, r = a - b", "build_qla.pl:1421-1456, QLA_c_eq_c_minus_c"),
    "r_times_plus": ("r = a[0] b + c, real a", "build_qla.pl:1837-1887, QLA_c_eq_r_times_c_plus_c"),
    "r_times_minus": ("r = a[0] b - c, real a", "build_qla.pl:1837-1887, QLA_c_eq_r_times_c_minus_c"),
    "c_times_plus": ("r = a[0] b + c, complex a", "build_qla.pl:1889-1935, QLA_c_eq_c_times_c_plus_c"),
    "c_times_minus": ("r = a[0] b - c, complex a", "build_qla.pl:1889-1935, QLA_c_eq_c_times_c_minus_c"),
    "zero": ("zero fill, r = 0", "build_qla.pl:2561-2580, QLA_c_eq_r(e, 0.)"),
    "rec": ("constant fill, r eqop a[0] (one record for every site)", "build_qla.pl:2584-2603, QLA_c_<eqop>_c"),
    "norm2": ("site-local squared norm, R eqop sum of |a|^2 over the record (double sum, operands promoted)",
              "build_qla.pl:495-523, make_code_binary.pl:38-97 (make_code_binary_dot)"),
    "dot": ("site-local inner product, C eqop sum of conj(a) b over the record (double sum)",
            "build_qla.pl:1758-1783, make_code_binary.pl:38-97"),
    "redot": ("real part of the site-local inner product, R eqop Re sum conj(a) b",
              "build_qla.pl:1810-1831, make_code_binary.pl:38-97"),
    "re_trace": ("real part of the trace, R = Re tr a", "build_qla.pl:933-981, expressions_tensor.pl:353-389"),
    "im_trace": ("imaginary part of the trace, R = Im tr a", "build_qla.pl:933-981, expressions_tensor.pl:353-389"),
    "trace": ("trace, C = tr a", "build_qla.pl:933-981, expressions_tensor.pl:353-389"),
    "antiherm": ("traceless anti-Hermitian part, M = (a - a^dagger)/2 - i Im tr(a)/3",
                 "build_qla.pl:983-1000, expressions_tensor.pl:138-232"),
    "fieldR": ("times a real field, r eqop a(x) b(x) (`Ta`: the adjoint of b)",
               "build_qla.pl:1517-1548, QLA_c_<eqop>_r_times_{c,ca}"),
    "fieldC": ("times a complex field, r eqop a(x) b(x) (`Ca`: conj(a); `Ta`: the adjoint of b)",
               "build_qla.pl:1551-1586, QLA_c_<eqop>_{c,ca}_times_{c,ca} (include/qla_complex.h:326-403)"),
    "int": ("integer fields (QLA_Int): copies, the int constant, sums, division, products, shifts, mod, "
            "max, min, comparisons, or / and / xor / not, copymask and the fills; C integer semantics",
            "build_qla.pl:356-402, 1256-1294, 1421-1516, 1837-1888, 1936-2088, 2561-2620"),
    "mask": ("copymask, r = a where the int field b is nonzero (other records untouched)",
             "build_qla.pl:2058-2088, expressions_tensor.pl:1065-1070"),
    "cmp": ("comparisons of real fields, the int r = (a op b) for eq ne gt lt ge le",
            "build_qla.pl:1976-2018, expressions_tensor.pl:1009-1019"),
    "sfn": ("elementary functions of real and complex fields (eq): R = f(R) for cos ... ceil, "
            "R = R op R for mod max min pow atan2, R = R ldexp I, C = cexpi R, R = norm / arg C, "
            "C = cexp / csqrt / clog C; transcendental results within a few units in the last place of the "
            "C library's", "build_qla.pl:194-355, make_code_unary.pl:103-176, expressions_tensor.pl:1021-1075"),
    "sconv": ("real / complex conversions C = R, C = R + iR, R = re C, R = im C and the division of "
              "real and complex fields (eq; C / C as QLA_c_eq_c_div_c)",
              "build_qla.pl:623-708, 1457-1482, include/qla_complex.h:252-255, 417-427"),
    "comp_elem": ("element access, C = a(ic, ...) and the insertion r(ic, ...) = a (eq; the indices "
                  "the type has: V ic; H, D ic, is; M ic, jc; P ic, is, jc, js; out of range: an error, "
                  "nothing written)", "build_qla.pl:767-820, expressions_tensor.pl:297-347"),
    "comp_colorvec": ("colour-column access, V = a(., fixed) and its insertion (M: jc; H, D: is; "
                      "P: is, jc, js)", "build_qla.pl:822-876, expressions_tensor.pl:297-347"),
    "comp_diracvec": ("Dirac-column access, D = P(., ., jc, js) and its insertion",
                      "build_qla.pl:878-932, expressions_tensor.pl:297-347"),
    "comp_spintrace": ("spin trace, M(ic, jc) = sum over is of P(ic, is, jc, is) from 0 in is order",
                       "build_qla.pl:1066-1086, expressions_tensor.pl:394-430"),
    "pmul0": ("propagator from a colour matrix times a propagator, r eqop a b (either factor's adjoint: "
              "`Ma`, `Pa`; the matrix acts on colour only)",
              "build_qla.pl:1587-1620, expressions_tensor.pl:544-630"),
    "pmul1": ("propagator from a propagator times a colour matrix, r eqop a b (`Pa`, `Ma`: adjoints)",
              "build_qla.pl:1622-1662, expressions_tensor.pl:544-630"),
    "pmul2": ("propagator from a product of propagators, r eqop a b as 12 x 12 matrices over (colour, spin) "
              "(`Pa`: adjoints)", "build_qla.pl:1622-1662, expressions_tensor.pl:544-630"),
    "outer": ("colour matrix from the outer product, r eqop a b^dagger",
              "build_qla.pl:1730-1756, QLA_c_<eqop>_c_times_ca"),
    "sum": ("global sum of a field, r = sum over sites of a (double sum per real column; DF: float fields, "
            "double result)", "build_qla.pl:2449-2521, make_code_unary.pl:318-366"),
    "gamma_left": ("left multiplication by Gamma_mu = g0^b0 g1^b1 g2^b2 g3^b3 (bits b of mu = 0..15), r = Gamma_mu a",
                   "build_qla.pl:1356-1384, expressions_gamma.pl:123-198, 387-470"),
    "gamma_right": ("right multiplication by Gamma_mu, r = a Gamma_mu", "build_qla.pl:1386-1415, expressions_gamma.pl"),
    "diag": ("diagonal fill, r eqop a[0] times the unit matrix",
             "build_qla.pl:2605-2617, expressions_tensor.pl:77-128 (QLA_c_<eqop>_r(e, 0.), then the diagonal)"),
}


def _vec_header_variants() -> list:
    """the families declared in qla_mi355x_vecops.h, in header order"""
    return (vecops() + fills() + gammas() + locals_() + sums() + fieldmuls() + pmuls() + comps() + sconvs() + sfns()
            + masks() + ints())


def vecops_header_text() -> str:
    lines = [
        "/*",
        " * qla_mi355x_vecops.h -- field linear algebra of the QLA C ABI on MI355X:",
        " * copies and increments, Hermitian conjugate, transpose, complex conjugate,",
        " * times i, times a real or complex constant, sums, differences and",
        " * a*b +- c with a constant a, the fills, the gamma-matrix multiplications and",
        " * the site-local norms, inner products, traces and anti-Hermitian parts, for",
        " * the types V, H, D, M, P.",
        " * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        " * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device",
        " * semantics are documented there).",
        " *",
        " * Names and argument order are QLA's (lib/perl/prototype.pl:450-688): the",
        " * constant `a` is a pointer to one QLA_F_Real / QLA_F_Complex (D: double)",
        " * in host or device memory; a host value is read during the call, a",
        " * device one when the kernel runs.  A `p` operand is a pointer array;",
        " * x-forms use every operand at index[i]; `x` operands use their own index",
        " * arrays (QLA_INDEX_ALL forms).",
        " */",
        "#ifndef QLA_MI355X_VECOPS_H",
        "#define QLA_MI355X_VECOPS_H",
    ]
    last = None
    for v in _vec_header_variants():
        if isinstance(v, IntVariant):
            key = ("I", "int")
        elif isinstance(v, MaskVariant):
            key = (v.prec[0] if v.t in ("cmp", "R", "C") else v.prec, "cmp" if v.t == "cmp" else "mask")
        elif isinstance(v, SfnVariant):
            key = (v.prec[0], "sfn")
        elif isinstance(v, SconvVariant):
            key = (v.prec[0], "sconv")
        elif isinstance(v, CompVariant):
            key = (v.prec, "comp_" + v.qual)
        elif isinstance(v, PmulVariant):
            key = (v.prec, "pmul" + str(v.kind))
        elif isinstance(v, FieldVariant):
            key = (v.prec, "outer" if v.kind == "outer" else "field" + v.kind)
        elif isinstance(v, SumVariant):
            key = (v.lib, "sum")
        elif isinstance(v, GammaVariant):
            key = (v.prec, "gamma_right" if v.right else "gamma_left")
        elif isinstance(v, LocalVariant):
            key = (v.prec, v.kind)
        else:
            key = (v.prec, v.fam if isinstance(v, VecVariant) else v.kind)
        if key != last:
            what, ref = _VEC_DOC[key[1]]
            lines += ["", f"/* {key[0]}: {what}.", f" * Reference: {ref}. */"]
            last = key
        lines.append(_decl(v) + ";")
    lines += ["", "#endif", ""]
    return "\n".join(lines)


def vecops_inc_text() -> str:
    lines = [
        "// qla_vecops_abi.inc -- definitions of the field linear algebra of the C ABI.",
        "// GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        "// Included by qla_capi.hip inside extern \"C\"; vec_entry, fill_entry, gamma_entry, local_entry,",
        "// sum_entry, field_entry, outer_entry, pmul_entry, comp_entry, sconv_entry, sfn_entry, mask_entry and int_entry\n"
        "// are there.",
    ]
    for v in _vec_header_variants():
        lines += [_decl(v), "{", "    " + v.body(), "}"]
    lines.append("")
    return "\n".join(lines)


# generic names (prototype.pl:556-575, make_define_map :423-441, selection
# :216-260): for a library QLA_<P><C>, the fully generic QLA_<rest> when
# QLA_Precision == P and QLA_Colors == C, the colour-generic QLA_<P>_<rest>
# when QLA_Colors == C and the precision-generic QLA_<C>_<rest> when
# QLA_Precision == P; the mixed-precision libraries only get colour-generic
# names; N-colour names inject QLA_Nc as the first argument.
_GENERIC_LIBS = {
    "F3": ("'F'", "3"), "D3": ("'D'", "3"), "FN": ("'F'", "'N'"), "DF3": (None, "3"), "FD3": (None, "3"),
    "F": ("'F'", None), "D": ("'D'", None), "DF": (None, None),
}


def _split(name: str) -> tuple[str, str]:
    body = name[len("QLA_"):]
    lib, rest = body.split("_", 1)
    return lib, rest


def generic_header_text() -> str:
    # the integer-field functions carry no library letter: their names are already generic
    names = [v.name for v in all_symbols() if not isinstance(v, IntVariant)]
    groups: dict[tuple[str, str], list[str]] = {}
    for nm in names:
        lib, rest = _split(nm)
        prec, col = _GENERIC_LIBS[lib]
        p_letter, c_letter = lib[:-1], lib[-1]
        nc = lib == "FN"

        def mapping(generic: str) -> str:
            if nc:
                return f"#define {generic}(...) {nm}(QLA_Nc, __VA_ARGS__)"
            return f"#define {generic} {nm}"
        if col is None:
            # colourless libraries (QLA_F, QLA_D): only the fully generic name
            if prec is not None:
                groups.setdefault((f"QLA_Precision == {prec}", "precision-only"), []).append(mapping(f"QLA_{rest}"))
            continue
        if prec is not None:
            groups.setdefault((f"QLA_Precision == {prec} && QLA_Colors == {col}", "generic"), []).append(
                mapping(f"QLA_{rest}"))
            groups.setdefault((f"QLA_Precision == {prec}", "precision"), []).append(
                mapping(f"QLA_{c_letter}_{rest}"))
        groups.setdefault((f"QLA_Colors == {col}", "colour"), []).append(mapping(f"QLA_{p_letter}_{rest}"))
    lines = [
        "/*",
        " * qla_mi355x_generic.h -- QLA's generic function and type names for the",
        " * specific symbols of this library.",
        " * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.",
        " * Included by qla_mi355x.h.",
        " *",
        " * As in the reference (include/qla.h:8-100, lib/perl/prototype.pl:160-260,",
        " * 423-441, 556-575): QLA_Precision ('F' or 'D', default 'F'; or",
        " * QLA_PrecisionInt 1 / 2) and QLA_Colors (3 or 'N'; or QLA_Nc, default 3)",
        " * select the library behind the fully generic names (QLA_V_veq_M_times_V),",
        " * the colour-generic names (QLA_F_V_veq_M_times_V, QLA_DF_r_veq_norm2_D)",
        " * and the precision-generic names (QLA_3_V_veq_M_times_V).  N-colour names",
        " * pass QLA_Nc as the leading nc argument.",
        " */",
        "#ifndef QLA_MI355X_GENERIC_H",
        "#define QLA_MI355X_GENERIC_H",
        "",
        "#ifndef QLA_Precision",
        "#  if !defined(QLA_PrecisionInt) || QLA_PrecisionInt == 1",
        "#    define QLA_Precision 'F'",
        "#  elif QLA_PrecisionInt == 2",
        "#    define QLA_Precision 'D'",
        "#  else",
        "#    error \"qla_mi355x: QLA_PrecisionInt must be 1 (F) or 2 (D)\"",
        "#  endif",
        "#endif",
        "#ifndef QLA_Colors",
        "#  ifndef QLA_Nc",
        "#    define QLA_Nc 3",
        "#  endif",
        "#  if QLA_Nc == 1 || QLA_Nc == 2 || QLA_Nc == 3",
        "#    define QLA_Colors QLA_Nc",
        "#  elif QLA_Nc > 0",
        "#    define QLA_Colors 'N'",
        "#  else",
        "#    error \"qla_mi355x: bad QLA_Nc\"",
        "#  endif",
        "#elif !defined(QLA_Nc)",
        "#  if QLA_Colors == 'N'",
        "#    define QLA_Nc 3",
        "#  else",
        "#    define QLA_Nc QLA_Colors",
        "#  endif",
        "#endif",
        "",
        "/* generic types (include/qla_types.h:380-420 maps them the same way) */",
        "#ifndef _QLA_TYPES_H",
        "#if QLA_Precision == 'F'",
        "typedef QLA_F_Real QLA_Real;",
        "typedef QLA_F_Complex QLA_Complex;",
        "#elif QLA_Precision == 'D'",
        "typedef QLA_D_Real QLA_Real;",
        "typedef QLA_D_Complex QLA_Complex;",
        "#endif",
        "#if QLA_Colors == 3",
    ]
    for t in TYPE_NAMES.values():
        lines += [f"typedef QLA_F3_{t} QLA_F_{t};", f"typedef QLA_D3_{t} QLA_D_{t};"]
    lines += ["#if QLA_Precision == 'F'"]
    lines += [f"typedef QLA_F3_{t} QLA_{t};" for t in TYPE_NAMES.values()]
    lines += ["#elif QLA_Precision == 'D'"]
    lines += [f"typedef QLA_D3_{t} QLA_{t};" for t in TYPE_NAMES.values()]
    lines += ["#endif", "#endif", "#endif /* _QLA_TYPES_H */"]
    for (cond, _), defs in groups.items():
        lines += ["", f"#if {cond}"] + defs + ["#endif"]
    lines += ["", "#endif", ""]
    return "\n".join(lines)


_OUTPUTS = ((HEADER_OUT, header_text), (INC_OUT, inc_text), (FIELDS_HEADER_OUT, fields_header_text),
            (FIELDS_INC_OUT, fields_inc_text), (MATFUNCS_HEADER_OUT, matfuncs_header_text),
            (MATFUNCS_INC_OUT, matfuncs_inc_text), (VECOPS_HEADER_OUT, vecops_header_text),
            (VECOPS_INC_OUT, vecops_inc_text), (GENERIC_HEADER_OUT, generic_header_text))


def write() -> None:
    for path, text in _OUTPUTS:
        path.write_text(text())


def current() -> bool:
    """True when the committed outputs match the table."""
    return all(path.exists() and path.read_text() == text() for path, text in _OUTPUTS)


if __name__ == "__main__":
    write()
    print(f"wrote {', '.join(str(p) for p, _ in _OUTPUTS)}: {len(all_symbols())} symbols")
    sys.exit(0)
