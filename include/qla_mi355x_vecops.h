/*
 * qla_mi355x_vecops.h -- field linear algebra of the QLA C ABI on MI355X:
 * copies and increments, Hermitian conjugate, transpose, complex conjugate,
 * times i, times a real or complex constant, sums, differences and
 * a*b +- c with a constant a, the fills and the gamma-matrix multiplications,
 * for the types V, H, D, M, P.
 * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.
 * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device
 * semantics are documented there).
 *
 * Names and argument order are QLA's (lib/perl/prototype.pl:450-688): the
 * constant `a` is a pointer to one QLA_F_Real / QLA_F_Complex (D: double)
 * in host or device memory; a host value is read during the call, a
 * device one when the kernel runs.  A `p` operand is a pointer array;
 * x-forms use every operand at index[i]; `x` operands use their own index
 * arrays (QLA_INDEX_ALL forms).
 */
#ifndef QLA_MI355X_VECOPS_H
#define QLA_MI355X_VECOPS_H

/* F3: copy or increment, r eqop a; `Ta` is the Hermitian conjugate.
 * Reference: build_qla.pl:356-430, macros QLA_c_<eqop>_c / _ca (include/qla_complex.h:147-163). */
void QLA_F3_V_veq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_V_veq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_F3_V_xeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_V_xeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_F3_xV_eq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_V_eq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_F3_xV_eq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_V_vpeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_vpeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_V_xpeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_V_xpeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xV_peq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_peq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xV_peq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_V_vmeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_vmeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_V_xmeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_V_xmeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xV_meq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_meq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xV_meq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_V_veqm_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_veqm_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_V_xeqm_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_V_xeqm_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xV_eqm_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_eqm_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xV_eqm_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_H_veq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_H_veq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_F3_H_xeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_H_xeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_F3_xH_eq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_H_eq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_F3_xH_eq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_H_vpeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_vpeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_H_xpeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_xpeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xH_peq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_peq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xH_peq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_H_vmeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_vmeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_H_xmeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_xmeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xH_meq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_meq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xH_meq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_H_veqm_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_veqm_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_H_xeqm_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_xeqm_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xH_eqm_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_eqm_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xH_eqm_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_D_veq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_D_veq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_F3_D_xeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_D_xeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_F3_xD_eq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_D_eq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_F3_xD_eq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_D_vpeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_vpeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_D_xpeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_xpeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xD_peq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_peq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xD_peq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_D_vmeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_vmeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_D_xmeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_xmeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xD_meq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_meq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xD_meq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_D_veqm_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_veqm_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_D_xeqm_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_xeqm_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xD_eqm_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_eqm_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xD_eqm_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_M_veq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_M_veq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_F3_M_xeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_M_xeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_F3_xM_eq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_M_eq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_F3_xM_eq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_M_vpeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_vpeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_M_xpeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_xpeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xM_peq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_peq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xM_peq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_M_vmeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_vmeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_M_xmeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_xmeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xM_meq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_meq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xM_meq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_M_veqm_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_veqm_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_M_xeqm_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_xeqm_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xM_eqm_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_eqm_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xM_eqm_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_M_veq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_veq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_M_xeq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_xeq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xM_eq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_eq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xM_eq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_M_vpeq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_vpeq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_M_xpeq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_M_xpeq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xM_peq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_peq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xM_peq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_M_vmeq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_vmeq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_M_xmeq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_M_xmeq_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xM_meq_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_meq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xM_meq_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_M_veqm_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_veqm_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_M_xeqm_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_M_xeqm_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xM_eqm_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_eqm_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xM_eqm_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_P_veq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_P_veq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_F3_P_xeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_P_xeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_F3_xP_eq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_P_eq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_F3_xP_eq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_P_vpeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_vpeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_P_xpeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_xpeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xP_peq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_peq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xP_peq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_P_vmeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_vmeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_P_xmeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_xmeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xP_meq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_meq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xP_meq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_P_veqm_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_veqm_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_P_xeqm_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_xeqm_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xP_eqm_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_eqm_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xP_eqm_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_P_veq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_veq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_F3_P_xeq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_xeq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_F3_xP_eq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_eq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_F3_xP_eq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_P_vpeq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_vpeq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_P_xpeq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_P_xpeq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xP_peq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_peq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xP_peq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_P_vmeq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_vmeq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_P_xmeq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_P_xmeq_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xP_meq_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_meq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xP_meq_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_P_veqm_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_veqm_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_P_xeqm_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_P_xeqm_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xP_eqm_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_eqm_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xP_eqm_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);

/* F3: transpose, r eqop a^T (P: the (colour, spin) pairs swapped).
 * Reference: build_qla.pl:432-460, lib/perl/headers.pl:68-90. */
void QLA_F3_M_veq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_M_veq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_F3_M_xeq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_F3_M_xeq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_F3_xM_eq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_M_eq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_F3_xM_eq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_M_vpeq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_vpeq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_M_xpeq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_M_xpeq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xM_peq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_peq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xM_peq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_F3_M_vmeq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_vmeq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_M_xmeq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_M_xmeq_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xM_meq_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_meq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xM_meq_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_F3_M_veqm_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_veqm_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_M_xeqm_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_M_xeqm_transpose_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xM_eqm_transpose_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_M_eqm_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xM_eqm_transpose_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_F3_P_veq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_P_veq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_F3_P_xeq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_F3_P_xeq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_F3_xP_eq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_P_eq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_F3_xP_eq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_P_vpeq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_vpeq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_P_xpeq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_P_xpeq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xP_peq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_peq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xP_peq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_F3_P_vmeq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_vmeq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_P_xmeq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_P_xmeq_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xP_meq_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_meq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xP_meq_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_F3_P_veqm_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_veqm_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_P_xeqm_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_P_xeqm_transpose_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xP_eqm_transpose_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_P_eqm_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xP_eqm_transpose_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);

/* F3: complex conjugate, r eqop conj(a).
 * Reference: build_qla.pl:462-493, QLA_c_<eqop>_ca. */
void QLA_F3_V_veq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_V_veq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_V_xeq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_V_xeq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xV_eq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_V_eq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xV_eq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_V_vpeq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_vpeq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_V_xpeq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_V_xpeq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xV_peq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_peq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xV_peq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_V_vmeq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_vmeq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_V_xmeq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_V_xmeq_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xV_meq_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_meq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xV_meq_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_V_veqm_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_veqm_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_V_xeqm_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_V_xeqm_conj_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xV_eqm_conj_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_V_eqm_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xV_eqm_conj_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_H_veq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_H_veq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_H_xeq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_H_xeq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xH_eq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_H_eq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xH_eq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_H_vpeq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_vpeq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_H_xpeq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_H_xpeq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xH_peq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_peq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xH_peq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_H_vmeq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_vmeq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_H_xmeq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_H_xmeq_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xH_meq_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_meq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xH_meq_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_H_veqm_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_veqm_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_H_xeqm_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_H_xeqm_conj_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xH_eqm_conj_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_H_eqm_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xH_eqm_conj_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_D_veq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_D_veq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_D_xeq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_D_xeq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xD_eq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_D_eq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xD_eq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_D_vpeq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_vpeq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_D_xpeq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_D_xpeq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xD_peq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_peq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xD_peq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_D_vmeq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_vmeq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_D_xmeq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_D_xmeq_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xD_meq_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_meq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xD_meq_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_D_veqm_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_veqm_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_D_xeqm_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_D_xeqm_conj_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xD_eqm_conj_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_D_eqm_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xD_eqm_conj_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_M_veq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_veq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_M_xeq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_M_xeq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xM_eq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_eq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xM_eq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_M_vpeq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_vpeq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_M_xpeq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_M_xpeq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xM_peq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_peq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xM_peq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_M_vmeq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_vmeq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_M_xmeq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_M_xmeq_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xM_meq_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_meq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xM_meq_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_M_veqm_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_veqm_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_M_xeqm_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_M_xeqm_conj_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xM_eqm_conj_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_eqm_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xM_eqm_conj_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_P_veq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_P_veq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_P_xeq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_P_xeq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xP_eq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_P_eq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xP_eq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_P_vpeq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_vpeq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_P_xpeq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_P_xpeq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xP_peq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_peq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xP_peq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_P_vmeq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_vmeq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_P_xmeq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_P_xmeq_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xP_meq_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_meq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xP_meq_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_P_veqm_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_veqm_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_P_xeqm_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_P_xeqm_conj_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xP_eqm_conj_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_P_eqm_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xP_eqm_conj_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);

/* F3: times i, r eqop i a.
 * Reference: build_qla.pl:1328-1354, QLA_c_<eqop>_ic (include/qla_complex.h:165-172). */
void QLA_F3_V_veq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_V_veq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_V_xeq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_V_xeq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xV_eq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_V_eq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xV_eq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_V_vpeq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_vpeq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_V_xpeq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_V_xpeq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xV_peq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_peq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xV_peq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_V_vmeq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_vmeq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_V_xmeq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_V_xmeq_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xV_meq_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_meq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xV_meq_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_V_veqm_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_veqm_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_V_xeqm_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_V_xeqm_i_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xV_eqm_i_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_V_eqm_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xV_eqm_i_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_H_veq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_H_veq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_H_xeq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_H_xeq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xH_eq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_H_eq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xH_eq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_H_vpeq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_vpeq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_H_xpeq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_H_xpeq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xH_peq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_peq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xH_peq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_H_vmeq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_vmeq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_H_xmeq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_H_xmeq_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xH_meq_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_meq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xH_meq_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_H_veqm_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_veqm_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_H_xeqm_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_H_xeqm_i_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xH_eqm_i_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_H_eqm_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xH_eqm_i_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_D_veq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_D_veq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_D_xeq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_D_xeq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xD_eq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_D_eq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xD_eq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_D_vpeq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_vpeq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_D_xpeq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_D_xpeq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xD_peq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_peq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xD_peq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_D_vmeq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_vmeq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_D_xmeq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_D_xmeq_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xD_meq_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_meq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xD_meq_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_D_veqm_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_veqm_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_D_xeqm_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_D_xeqm_i_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xD_eqm_i_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_D_eqm_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xD_eqm_i_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_M_veq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_veq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_M_xeq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_M_xeq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xM_eq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_M_eq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xM_eq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_M_vpeq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_vpeq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_M_xpeq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_M_xpeq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xM_peq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_peq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xM_peq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_M_vmeq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_vmeq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_M_xmeq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_M_xmeq_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xM_meq_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_meq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xM_meq_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_M_veqm_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_veqm_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_M_xeqm_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_M_xeqm_i_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xM_eqm_i_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_M_eqm_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xM_eqm_i_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_P_veq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_veq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_F3_P_xeq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_F3_P_xeq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_F3_xP_eq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_F3_P_eq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_F3_xP_eq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_P_vpeq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_vpeq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_P_xpeq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_P_xpeq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xP_peq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_peq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xP_peq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_P_vmeq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_vmeq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_P_xmeq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_P_xmeq_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xP_meq_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_meq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xP_meq_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_P_veqm_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_veqm_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_F3_P_xeqm_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_F3_P_xeqm_i_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_F3_xP_eqm_i_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_F3_P_eqm_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_F3_xP_eqm_i_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);

/* F3: times a real constant, r eqop a[0] b (`Ta`: the adjoint of b).
 * Reference: build_qla.pl:1256-1293, QLA_c_<eqop>_r_times_c / _ca (include/qla_complex.h:259-284). */
void QLA_F3_V_veq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_veq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_V_xeq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_V_xeq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xV_eq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_eq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xV_eq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_V_vpeq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_vpeq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xpeq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xpeq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_peq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_peq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_peq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_V_vmeq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_vmeq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xmeq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xmeq_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_meq_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_meq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_meq_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_V_veqm_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_veqm_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xeqm_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xeqm_r_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_eqm_r_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_eqm_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_eqm_r_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_veq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_veq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_H_xeq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_H_xeq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xH_eq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_eq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xH_eq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_H_vpeq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_vpeq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xpeq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xpeq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_peq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_peq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_peq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_vmeq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_vmeq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xmeq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xmeq_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_meq_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_meq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_meq_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_veqm_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_veqm_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xeqm_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xeqm_r_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_eqm_r_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_eqm_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_eqm_r_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_veq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_veq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_D_xeq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_D_xeq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xD_eq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_eq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xD_eq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_D_vpeq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_vpeq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xpeq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xpeq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_peq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_peq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_peq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_vmeq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_vmeq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xmeq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xmeq_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_meq_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_meq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_meq_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_veqm_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_veqm_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xeqm_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xeqm_r_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_eqm_r_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_eqm_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_eqm_r_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_veq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_M_xeq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_M_xeq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xM_eq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_eq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xM_eq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_M_vpeq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_vpeq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xpeq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xpeq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_peq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_peq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_peq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_vmeq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_vmeq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xmeq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xmeq_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_meq_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_meq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_meq_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veqm_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_veqm_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xeqm_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xeqm_r_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_eqm_r_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_eqm_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eqm_r_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_veq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xeq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xeq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_eq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_eq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_vpeq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_vpeq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xpeq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xpeq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_peq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_peq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_peq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_M_vmeq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_vmeq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xmeq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xmeq_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_meq_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_meq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_meq_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_M_veqm_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_veqm_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xeqm_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xeqm_r_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_eqm_r_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_eqm_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_eqm_r_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_veq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_veq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_P_xeq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_P_xeq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xP_eq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_eq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Real *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xP_eq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_P_vpeq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_vpeq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xpeq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xpeq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_peq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_peq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_peq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_vmeq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_vmeq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xmeq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xmeq_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_meq_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_meq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_meq_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_veqm_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_veqm_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xeqm_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xeqm_r_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_eqm_r_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_eqm_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eqm_r_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_veq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_veq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xeq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xeq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_eq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_eq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Real *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_vpeq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_vpeq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xpeq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xpeq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_peq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_peq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_peq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_vmeq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_vmeq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xmeq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xmeq_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_meq_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_meq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_meq_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_veqm_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_veqm_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xeqm_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xeqm_r_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_eqm_r_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_eqm_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Real *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_eqm_r_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Real *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);

/* F3: times a complex constant, r eqop a[0] b (`Ta`: the adjoint of b).
 * Reference: build_qla.pl:1295-1326, QLA_c_<eqop>_c_times_c / _ca (include/qla_complex.h:326-379). */
void QLA_F3_V_veq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_veq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_V_xeq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_V_xeq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xV_eq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_eq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xV_eq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_V_vpeq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_vpeq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xpeq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xpeq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_peq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_peq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_peq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_V_vmeq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_vmeq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xmeq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xmeq_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_meq_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_meq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_meq_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_V_veqm_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_veqm_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xeqm_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xeqm_c_times_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xV_eqm_c_times_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_eqm_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_eqm_c_times_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_veq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_veq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_H_xeq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_H_xeq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xH_eq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_eq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xH_eq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_H_vpeq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_vpeq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xpeq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xpeq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_peq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_peq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_peq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_vmeq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_vmeq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xmeq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xmeq_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_meq_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_meq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_meq_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_veqm_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_veqm_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xeqm_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xeqm_c_times_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xH_eqm_c_times_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_eqm_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_eqm_c_times_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_veq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_veq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_D_xeq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_D_xeq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xD_eq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_eq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xD_eq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_D_vpeq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_vpeq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xpeq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xpeq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_peq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_peq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_peq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_vmeq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_vmeq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xmeq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xmeq_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_meq_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_meq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_meq_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_veqm_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_veqm_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xeqm_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xeqm_c_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xD_eqm_c_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_eqm_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_eqm_c_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_veq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_M_xeq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_M_xeq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xM_eq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_eq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xM_eq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_M_vpeq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_vpeq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xpeq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xpeq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_peq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_peq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_peq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_vmeq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_vmeq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xmeq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xmeq_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_meq_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_meq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_meq_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veqm_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_veqm_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xeqm_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xeqm_c_times_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_eqm_c_times_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_eqm_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eqm_c_times_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_veq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xeq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xeq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xM_eq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_eq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_vpeq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_vpeq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xpeq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xpeq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_peq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_peq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_peq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_M_vmeq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_vmeq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xmeq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xmeq_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_meq_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_meq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_meq_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_M_veqm_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_veqm_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_xeqm_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xeqm_c_times_pMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xM_eqm_c_times_Ma(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_eqm_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xM_eqm_c_times_xMa(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_veq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_veq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_P_xeq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_P_xeq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_xP_eq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_eq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F_Complex *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xP_eq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_P_vpeq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_vpeq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xpeq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xpeq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_peq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_peq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_peq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_vmeq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_vmeq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xmeq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xmeq_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_meq_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_meq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_meq_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_veqm_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_veqm_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xeqm_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xeqm_c_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_eqm_c_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_eqm_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eqm_c_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_veq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_veq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xeq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xeq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_xP_eq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_eq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F_Complex *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_vpeq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_vpeq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xpeq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xpeq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_peq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_peq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_peq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_vmeq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_vmeq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xmeq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xmeq_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_meq_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_meq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_meq_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_P_veqm_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_veqm_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_P_xeqm_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_P_xeqm_c_times_pPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_xP_eqm_c_times_Pa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_P_eqm_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F_Complex *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_xP_eqm_c_times_xPa(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F_Complex *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);

/* F3: sum, r = a + b.
 * Reference: build_qla.pl:1421-1456, QLA_c_eq_c_plus_c. */
void QLA_F3_V_veq_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_V_veq_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_V_veq_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_veq_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_V_xeq_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_V_xeq_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_V_xeq_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_V_xeq_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_V_eq_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_V_eq_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_V_eq_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xV_eq_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_xV_eq_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xV_eq_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xV_eq_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_H_veq_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_H_veq_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_H_veq_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_veq_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_H_xeq_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_H_xeq_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_H_xeq_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_H_xeq_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_H_eq_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_H_eq_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_H_eq_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xH_eq_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_xH_eq_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xH_eq_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xH_eq_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_D_veq_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_D_veq_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_D_veq_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_veq_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_D_xeq_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_D_xeq_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_D_xeq_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_D_xeq_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_D_eq_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_D_eq_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_D_eq_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xD_eq_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_xD_eq_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xD_eq_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xD_eq_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_M_veq_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_M_veq_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_M_veq_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_veq_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_M_xeq_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_M_xeq_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_M_xeq_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_M_xeq_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_M_eq_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_M_eq_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_M_eq_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xM_eq_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_xM_eq_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xM_eq_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xM_eq_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_P_veq_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_P_veq_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_P_veq_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_veq_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_P_xeq_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_P_xeq_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_P_xeq_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_P_xeq_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_P_eq_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_P_eq_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_P_eq_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xP_eq_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_xP_eq_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_xP_eq_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xP_eq_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);

/* F3: difference, r = a - b.
 * Reference: build_qla.pl:1421-1456, QLA_c_eq_c_minus_c. */
void QLA_F3_V_veq_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_veq_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_V_veq_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_V_veq_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_V_xeq_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_V_xeq_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_V_xeq_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_V_xeq_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_V_eq_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_V_eq_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_V_eq_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_eq_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xV_eq_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xV_eq_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_xV_eq_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_H_veq_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_veq_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_H_veq_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_H_veq_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_H_xeq_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_H_xeq_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_H_xeq_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_H_xeq_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_H_eq_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_H_eq_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_H_eq_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_eq_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xH_eq_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xH_eq_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_xH_eq_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_D_veq_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_veq_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_D_veq_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_D_veq_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_D_xeq_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_D_xeq_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_D_xeq_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_D_xeq_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_D_eq_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_D_eq_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_D_eq_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_eq_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xD_eq_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xD_eq_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_xD_eq_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_veq_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_veq_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_M_veq_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_M_veq_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_M_xeq_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_M_xeq_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_M_xeq_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_M_xeq_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_M_eq_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_M_eq_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_M_eq_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eq_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xM_eq_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xM_eq_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_xM_eq_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_P_veq_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_veq_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_F3_P_veq_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_P_veq_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_P_xeq_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_F3_P_xeq_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_F3_P_xeq_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_P_xeq_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_P_eq_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_F3_P_eq_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_P_eq_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eq_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_F3_xP_eq_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_xP_eq_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_xP_eq_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);

/* F3: r = a[0] b + c, real a.
 * Reference: build_qla.pl:1837-1887, QLA_c_eq_r_times_c_plus_c. */
void QLA_F3_V_veq_r_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_V_veq_r_times_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_V_veq_r_times_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_veq_r_times_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_V_xeq_r_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_V_xeq_r_times_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_r_times_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_r_times_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_V_eq_r_times_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_V_eq_r_times_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_V_eq_r_times_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xV_eq_r_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xV_eq_r_times_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xV_eq_r_times_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xV_eq_r_times_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_H_veq_r_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_H_veq_r_times_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_H_veq_r_times_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_veq_r_times_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_H_xeq_r_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_H_xeq_r_times_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_r_times_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_r_times_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_H_eq_r_times_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_H_eq_r_times_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_H_eq_r_times_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xH_eq_r_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xH_eq_r_times_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xH_eq_r_times_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xH_eq_r_times_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_D_veq_r_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_D_veq_r_times_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_D_veq_r_times_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_veq_r_times_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_D_xeq_r_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_D_xeq_r_times_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_r_times_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_r_times_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_D_eq_r_times_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_D_eq_r_times_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_D_eq_r_times_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xD_eq_r_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xD_eq_r_times_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xD_eq_r_times_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xD_eq_r_times_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_M_veq_r_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_M_veq_r_times_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_M_veq_r_times_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_veq_r_times_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_M_xeq_r_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_M_xeq_r_times_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_r_times_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_r_times_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_M_eq_r_times_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_M_eq_r_times_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_M_eq_r_times_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xM_eq_r_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xM_eq_r_times_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xM_eq_r_times_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xM_eq_r_times_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_P_veq_r_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_P_veq_r_times_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_P_veq_r_times_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_veq_r_times_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_P_xeq_r_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_P_xeq_r_times_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_r_times_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_r_times_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_P_eq_r_times_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_P_eq_r_times_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_P_eq_r_times_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xP_eq_r_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Real *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xP_eq_r_times_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xP_eq_r_times_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xP_eq_r_times_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);

/* F3: r = a[0] b - c, real a.
 * Reference: build_qla.pl:1837-1887, QLA_c_eq_r_times_c_minus_c. */
void QLA_F3_V_veq_r_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_veq_r_times_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_V_veq_r_times_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_V_veq_r_times_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_V_xeq_r_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_r_times_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_V_xeq_r_times_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_V_xeq_r_times_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_V_eq_r_times_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_V_eq_r_times_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_eq_r_times_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xV_eq_r_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xV_eq_r_times_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xV_eq_r_times_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xV_eq_r_times_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_H_veq_r_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_veq_r_times_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_H_veq_r_times_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_H_veq_r_times_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_H_xeq_r_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_r_times_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_H_xeq_r_times_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_H_xeq_r_times_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_H_eq_r_times_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_H_eq_r_times_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_eq_r_times_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xH_eq_r_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xH_eq_r_times_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xH_eq_r_times_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xH_eq_r_times_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_D_veq_r_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_veq_r_times_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_D_veq_r_times_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_D_veq_r_times_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_D_xeq_r_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_r_times_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_D_xeq_r_times_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_D_xeq_r_times_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_D_eq_r_times_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_D_eq_r_times_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_eq_r_times_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xD_eq_r_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xD_eq_r_times_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xD_eq_r_times_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xD_eq_r_times_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_M_veq_r_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_veq_r_times_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_M_veq_r_times_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_M_veq_r_times_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_M_xeq_r_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_r_times_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_M_xeq_r_times_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_M_xeq_r_times_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_M_eq_r_times_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_M_eq_r_times_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_eq_r_times_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xM_eq_r_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xM_eq_r_times_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xM_eq_r_times_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xM_eq_r_times_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_P_veq_r_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_veq_r_times_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_P_veq_r_times_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_P_veq_r_times_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_P_xeq_r_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_r_times_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_P_xeq_r_times_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_P_xeq_r_times_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_P_eq_r_times_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_P_eq_r_times_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_eq_r_times_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xP_eq_r_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Real *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xP_eq_r_times_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xP_eq_r_times_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Real *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xP_eq_r_times_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Real *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);

/* F3: r = a[0] b + c, complex a.
 * Reference: build_qla.pl:1889-1935, QLA_c_eq_c_times_c_plus_c. */
void QLA_F3_V_veq_c_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_V_veq_c_times_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_V_veq_c_times_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_veq_c_times_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_V_xeq_c_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_V_xeq_c_times_V_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_c_times_pV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_c_times_pV_plus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_V_eq_c_times_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_V_eq_c_times_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_V_eq_c_times_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xV_eq_c_times_V_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xV_eq_c_times_V_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xV_eq_c_times_xV_plus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xV_eq_c_times_xV_plus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_H_veq_c_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_H_veq_c_times_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_H_veq_c_times_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_veq_c_times_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_H_xeq_c_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_H_xeq_c_times_H_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_c_times_pH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_c_times_pH_plus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_H_eq_c_times_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_H_eq_c_times_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_H_eq_c_times_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xH_eq_c_times_H_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xH_eq_c_times_H_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xH_eq_c_times_xH_plus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xH_eq_c_times_xH_plus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_D_veq_c_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_D_veq_c_times_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_D_veq_c_times_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_veq_c_times_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_D_xeq_c_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_D_xeq_c_times_D_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_c_times_pD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_c_times_pD_plus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_D_eq_c_times_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_D_eq_c_times_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_D_eq_c_times_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xD_eq_c_times_D_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xD_eq_c_times_D_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xD_eq_c_times_xD_plus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xD_eq_c_times_xD_plus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_M_veq_c_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_M_veq_c_times_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_M_veq_c_times_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_veq_c_times_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_M_xeq_c_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_M_xeq_c_times_M_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_c_times_pM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_c_times_pM_plus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_M_eq_c_times_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_M_eq_c_times_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_M_eq_c_times_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xM_eq_c_times_M_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xM_eq_c_times_M_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xM_eq_c_times_xM_plus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xM_eq_c_times_xM_plus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_P_veq_c_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_P_veq_c_times_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_F3_P_veq_c_times_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_veq_c_times_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_P_xeq_c_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_F3_P_xeq_c_times_P_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_c_times_pP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_c_times_pP_plus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_P_eq_c_times_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_F3_P_eq_c_times_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_P_eq_c_times_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xP_eq_c_times_P_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_F3_xP_eq_c_times_P_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_xP_eq_c_times_xP_plus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xP_eq_c_times_xP_plus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);

/* F3: r = a[0] b - c, complex a.
 * Reference: build_qla.pl:1889-1935, QLA_c_eq_c_times_c_minus_c. */
void QLA_F3_V_veq_c_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_veq_c_times_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_V_veq_c_times_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_V_veq_c_times_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_V_xeq_c_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_V_xeq_c_times_V_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_V_xeq_c_times_pV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_V_xeq_c_times_pV_minus_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_V_eq_c_times_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_V_eq_c_times_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_V_eq_c_times_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xV_eq_c_times_V_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xV_eq_c_times_V_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xV_eq_c_times_xV_minus_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xV_eq_c_times_xV_minus_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_H_veq_c_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_veq_c_times_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_H_veq_c_times_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_H_veq_c_times_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_H_xeq_c_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_H_xeq_c_times_H_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_H_xeq_c_times_pH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_H_xeq_c_times_pH_minus_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_H_eq_c_times_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_H_eq_c_times_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_H_eq_c_times_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xH_eq_c_times_H_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xH_eq_c_times_H_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xH_eq_c_times_xH_minus_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xH_eq_c_times_xH_minus_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_D_veq_c_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_veq_c_times_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_D_veq_c_times_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_D_veq_c_times_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_D_xeq_c_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_D_xeq_c_times_D_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_D_xeq_c_times_pD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_D_xeq_c_times_pD_minus_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_D_eq_c_times_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_D_eq_c_times_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_D_eq_c_times_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xD_eq_c_times_D_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xD_eq_c_times_D_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xD_eq_c_times_xD_minus_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xD_eq_c_times_xD_minus_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_M_veq_c_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_veq_c_times_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_M_veq_c_times_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_M_veq_c_times_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_M_xeq_c_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_M_xeq_c_times_M_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_M_xeq_c_times_pM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_M_xeq_c_times_pM_minus_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_M_eq_c_times_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_M_eq_c_times_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_M_eq_c_times_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xM_eq_c_times_M_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xM_eq_c_times_M_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xM_eq_c_times_xM_minus_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xM_eq_c_times_xM_minus_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_F3_P_veq_c_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_veq_c_times_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_F3_P_veq_c_times_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_P_veq_c_times_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_F3_P_xeq_c_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_F3_P_xeq_c_times_P_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_F3_P_xeq_c_times_pP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_F3_P_xeq_c_times_pP_minus_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_F3_P_eq_c_times_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_F3_P_eq_c_times_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_P_eq_c_times_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xP_eq_c_times_P_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_F3_xP_eq_c_times_P_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_F3_xP_eq_c_times_xP_minus_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_F3_xP_eq_c_times_xP_minus_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_F_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);

/* D3: copy or increment, r eqop a; `Ta` is the Hermitian conjugate.
 * Reference: build_qla.pl:356-430, macros QLA_c_<eqop>_c / _ca (include/qla_complex.h:147-163). */
void QLA_D3_V_veq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_V_veq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_D3_V_xeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_V_xeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_D3_xV_eq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_V_eq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_D3_xV_eq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_V_vpeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_vpeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_V_xpeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_V_xpeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xV_peq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_peq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xV_peq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_V_vmeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_vmeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_V_xmeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_V_xmeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xV_meq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_meq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xV_meq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_V_veqm_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_veqm_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_V_xeqm_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_V_xeqm_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xV_eqm_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_eqm_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xV_eqm_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_H_veq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_H_veq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_D3_H_xeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_H_xeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_D3_xH_eq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_H_eq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_D3_xH_eq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_H_vpeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_vpeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_H_xpeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_xpeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xH_peq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_peq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xH_peq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_H_vmeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_vmeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_H_xmeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_xmeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xH_meq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_meq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xH_meq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_H_veqm_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_veqm_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_H_xeqm_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_xeqm_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xH_eqm_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_eqm_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xH_eqm_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_D_veq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_D_veq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_D3_D_xeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_D_xeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_D3_xD_eq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_D_eq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_D3_xD_eq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_D_vpeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_vpeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_D_xpeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_xpeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xD_peq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_peq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xD_peq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_D_vmeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_vmeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_D_xmeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_xmeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xD_meq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_meq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xD_meq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_D_veqm_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_veqm_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_D_xeqm_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_xeqm_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xD_eqm_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_eqm_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xD_eqm_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_M_veq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_M_veq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_D3_M_xeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_M_xeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_D3_xM_eq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_M_eq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_D3_xM_eq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_M_vpeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_vpeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_M_xpeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_xpeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xM_peq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_peq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xM_peq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_M_vmeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_vmeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_M_xmeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_xmeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xM_meq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_meq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xM_meq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_M_veqm_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_veqm_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_M_xeqm_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_xeqm_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xM_eqm_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_eqm_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xM_eqm_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_M_veq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_veq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_M_xeq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_xeq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xM_eq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_eq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xM_eq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_M_vpeq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_vpeq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_M_xpeq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_M_xpeq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xM_peq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_peq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xM_peq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_M_vmeq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_vmeq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_M_xmeq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_M_xmeq_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xM_meq_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_meq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xM_meq_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_M_veqm_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_veqm_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_M_xeqm_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_M_xeqm_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xM_eqm_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_eqm_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xM_eqm_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_P_veq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_P_veq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                     int n);
void QLA_D3_P_xeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_P_xeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                     int *index,
                     int n);
void QLA_D3_xP_eq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    int *indexr,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_P_eq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *indexa,
                    int n);
void QLA_D3_xP_eq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_P_vpeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_vpeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_P_xpeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_xpeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xP_peq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_peq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xP_peq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_P_vmeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_vmeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_P_xmeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_xmeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xP_meq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_meq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xP_meq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_P_veqm_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_veqm_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_P_xeqm_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_xeqm_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xP_eqm_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_eqm_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xP_eqm_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_P_veq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_veq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_D3_P_xeq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_xeq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_D3_xP_eq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_eq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_D3_xP_eq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_P_vpeq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_vpeq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_P_xpeq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_P_xpeq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xP_peq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_peq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xP_peq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_P_vmeq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_vmeq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_P_xmeq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_P_xmeq_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xP_meq_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_meq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xP_meq_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_P_veqm_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_veqm_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_P_xeqm_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_P_xeqm_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xP_eqm_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_eqm_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xP_eqm_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);

/* D3: transpose, r eqop a^T (P: the (colour, spin) pairs swapped).
 * Reference: build_qla.pl:432-460, lib/perl/headers.pl:68-90. */
void QLA_D3_M_veq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_M_veq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_D3_M_xeq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_D3_M_xeq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_D3_xM_eq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_M_eq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_D3_xM_eq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_M_vpeq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_vpeq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_M_xpeq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_M_xpeq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xM_peq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_peq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xM_peq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_D3_M_vmeq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_vmeq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_M_xmeq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_M_xmeq_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xM_meq_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_meq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xM_meq_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_D3_M_veqm_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_veqm_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_M_xeqm_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_M_xeqm_transpose_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xM_eqm_transpose_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_M_eqm_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xM_eqm_transpose_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_D3_P_veq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_P_veq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_D3_P_xeq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_D3_P_xeq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_D3_xP_eq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_P_eq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_D3_xP_eq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_P_vpeq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_vpeq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_P_xpeq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_P_xpeq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xP_peq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_peq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xP_peq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_D3_P_vmeq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_vmeq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_P_xmeq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_P_xmeq_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xP_meq_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_meq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xP_meq_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);
void QLA_D3_P_veqm_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_veqm_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_P_xeqm_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_P_xeqm_transpose_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xP_eqm_transpose_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_P_eqm_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xP_eqm_transpose_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);

/* D3: complex conjugate, r eqop conj(a).
 * Reference: build_qla.pl:462-493, QLA_c_<eqop>_ca. */
void QLA_D3_V_veq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_V_veq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_V_xeq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_V_xeq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xV_eq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_V_eq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xV_eq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_V_vpeq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_vpeq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_V_xpeq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_V_xpeq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xV_peq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_peq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xV_peq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_V_vmeq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_vmeq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_V_xmeq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_V_xmeq_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xV_meq_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_meq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xV_meq_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_V_veqm_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_veqm_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_V_xeqm_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_V_xeqm_conj_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xV_eqm_conj_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_V_eqm_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xV_eqm_conj_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_H_veq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_H_veq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_H_xeq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_H_xeq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xH_eq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_H_eq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xH_eq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_H_vpeq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_vpeq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_H_xpeq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_H_xpeq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xH_peq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_peq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xH_peq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_H_vmeq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_vmeq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_H_xmeq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_H_xmeq_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xH_meq_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_meq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xH_meq_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_H_veqm_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_veqm_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_H_xeqm_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_H_xeqm_conj_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xH_eqm_conj_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_H_eqm_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xH_eqm_conj_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_D_veq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_D_veq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_D_xeq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_D_xeq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xD_eq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_D_eq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xD_eq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_D_vpeq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_vpeq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_D_xpeq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_D_xpeq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xD_peq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_peq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xD_peq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_D_vmeq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_vmeq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_D_xmeq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_D_xmeq_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xD_meq_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_meq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xD_meq_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_D_veqm_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_veqm_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_D_xeqm_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_D_xeqm_conj_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xD_eqm_conj_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_D_eqm_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xD_eqm_conj_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_M_veq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_veq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_M_xeq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_M_xeq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xM_eq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_eq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xM_eq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_M_vpeq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_vpeq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_M_xpeq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_M_xpeq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xM_peq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_peq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xM_peq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_M_vmeq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_vmeq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_M_xmeq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_M_xmeq_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xM_meq_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_meq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xM_meq_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_M_veqm_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_veqm_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_M_xeqm_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_M_xeqm_conj_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xM_eqm_conj_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_eqm_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xM_eqm_conj_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_P_veq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_P_veq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_P_xeq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_P_xeq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xP_eq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_P_eq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                         QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xP_eq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_P_vpeq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_vpeq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_P_xpeq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_P_xpeq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xP_peq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_peq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xP_peq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_P_vmeq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_vmeq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_P_xmeq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_P_xmeq_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xP_meq_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_meq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xP_meq_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_P_veqm_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_veqm_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_P_xeqm_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_P_xeqm_conj_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xP_eqm_conj_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_P_eqm_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xP_eqm_conj_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);

/* D3: times i, r eqop i a.
 * Reference: build_qla.pl:1328-1354, QLA_c_<eqop>_ic (include/qla_complex.h:165-172). */
void QLA_D3_V_veq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_V_veq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_V_xeq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_V_xeq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xV_eq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_V_eq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xV_eq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_V_vpeq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_vpeq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_V_xpeq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_V_xpeq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xV_peq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_peq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xV_peq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_V_vmeq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_vmeq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_V_xmeq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_V_xmeq_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xV_meq_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_meq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xV_meq_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_V_veqm_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_veqm_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_V_xeqm_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_V_xeqm_i_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xV_eqm_i_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_V_eqm_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xV_eqm_i_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_H_veq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_H_veq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_H_xeq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_H_xeq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xH_eq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_H_eq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xH_eq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_H_vpeq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_vpeq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_H_xpeq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_H_xpeq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xH_peq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_peq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xH_peq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_H_vmeq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_vmeq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_H_xmeq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_H_xmeq_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xH_meq_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_meq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xH_meq_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_H_veqm_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_veqm_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_H_xeqm_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_H_xeqm_i_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xH_eqm_i_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_H_eqm_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xH_eqm_i_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_D_veq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_D_veq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_D_xeq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_D_xeq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xD_eq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_D_eq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xD_eq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_D_vpeq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_vpeq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_D_xpeq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_D_xpeq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xD_peq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_peq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xD_peq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_D_vmeq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_vmeq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_D_xmeq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_D_xmeq_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xD_meq_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_meq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xD_meq_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_D_veqm_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_veqm_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_D_xeqm_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_D_xeqm_i_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xD_eqm_i_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_D_eqm_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xD_eqm_i_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_M_veq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_veq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_M_xeq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_M_xeq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xM_eq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_M_eq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xM_eq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_M_vpeq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_vpeq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_M_xpeq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_M_xpeq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xM_peq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_peq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xM_peq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_M_vmeq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_vmeq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_M_xmeq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_M_xmeq_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xM_meq_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_meq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xM_meq_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_M_veqm_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_veqm_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_M_xeqm_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_M_xeqm_i_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xM_eqm_i_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_M_eqm_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xM_eqm_i_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_P_veq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_veq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int n);
void QLA_D3_P_xeq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *index,
                      int n);
void QLA_D3_P_xeq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                       int *index,
                       int n);
void QLA_D3_xP_eq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int n);
void QLA_D3_P_eq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_D3_xP_eq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_P_vpeq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_vpeq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_P_xpeq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_P_xpeq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xP_peq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_peq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xP_peq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_P_vmeq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_vmeq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_P_xmeq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_P_xmeq_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xP_meq_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_meq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xP_meq_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_P_veqm_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_veqm_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int n);
void QLA_D3_P_xeqm_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *index,
                       int n);
void QLA_D3_P_xeqm_i_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                        int *index,
                        int n);
void QLA_D3_xP_eqm_i_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *indexr,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int n);
void QLA_D3_P_eqm_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                       int *indexa,
                       int n);
void QLA_D3_xP_eqm_i_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);

/* D3: times a real constant, r eqop a[0] b (`Ta`: the adjoint of b).
 * Reference: build_qla.pl:1256-1293, QLA_c_<eqop>_r_times_c / _ca (include/qla_complex.h:259-284). */
void QLA_D3_V_veq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_veq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_V_xeq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_V_xeq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xV_eq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_eq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xV_eq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_V_vpeq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_vpeq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xpeq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xpeq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_peq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_peq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_peq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_V_vmeq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_vmeq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xmeq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xmeq_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_meq_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_meq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_meq_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_V_veqm_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_veqm_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xeqm_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xeqm_r_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_eqm_r_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_eqm_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_eqm_r_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_veq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_veq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_H_xeq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_H_xeq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xH_eq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_eq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xH_eq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_H_vpeq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_vpeq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xpeq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xpeq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_peq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_peq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_peq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_vmeq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_vmeq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xmeq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xmeq_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_meq_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_meq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_meq_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_veqm_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_veqm_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xeqm_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xeqm_r_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_eqm_r_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_eqm_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_eqm_r_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_veq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_veq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_D_xeq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_D_xeq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xD_eq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_eq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xD_eq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_D_vpeq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_vpeq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xpeq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xpeq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_peq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_peq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_peq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_vmeq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_vmeq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xmeq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xmeq_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_meq_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_meq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_meq_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_veqm_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_veqm_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xeqm_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xeqm_r_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_eqm_r_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_eqm_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_eqm_r_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_veq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_M_xeq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_M_xeq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xM_eq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_eq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xM_eq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_M_vpeq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_vpeq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xpeq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xpeq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_peq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_peq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_peq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_vmeq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_vmeq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xmeq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xmeq_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_meq_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_meq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_meq_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veqm_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_veqm_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xeqm_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xeqm_r_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_eqm_r_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_eqm_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eqm_r_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_veq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xeq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xeq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_eq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_eq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_vpeq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_vpeq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xpeq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xpeq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_peq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_peq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_peq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_M_vmeq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_vmeq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xmeq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xmeq_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_meq_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_meq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_meq_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_M_veqm_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_veqm_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xeqm_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xeqm_r_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_eqm_r_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_eqm_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_eqm_r_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_veq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_veq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_P_xeq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_P_xeq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xP_eq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_eq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Real *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xP_eq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_P_vpeq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_vpeq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xpeq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xpeq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_peq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_peq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_peq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_vmeq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_vmeq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xmeq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xmeq_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_meq_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_meq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_meq_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_veqm_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_veqm_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xeqm_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xeqm_r_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_eqm_r_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_eqm_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eqm_r_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_veq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_veq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xeq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xeq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_eq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_eq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Real *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_vpeq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_vpeq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xpeq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xpeq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_peq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_peq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_peq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_vmeq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_vmeq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xmeq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xmeq_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_meq_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_meq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_meq_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_veqm_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_veqm_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xeqm_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xeqm_r_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_eqm_r_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_eqm_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Real *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_eqm_r_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Real *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);

/* D3: times a complex constant, r eqop a[0] b (`Ta`: the adjoint of b).
 * Reference: build_qla.pl:1295-1326, QLA_c_<eqop>_c_times_c / _ca (include/qla_complex.h:326-379). */
void QLA_D3_V_veq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_veq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_V_xeq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_V_xeq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xV_eq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_eq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xV_eq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_V_vpeq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_vpeq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xpeq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xpeq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_peq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_peq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_peq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_V_vmeq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_vmeq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xmeq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xmeq_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_meq_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_meq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_meq_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_V_veqm_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_veqm_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xeqm_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xeqm_c_times_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xV_eqm_c_times_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_eqm_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_eqm_c_times_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_veq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_veq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_H_xeq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_H_xeq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xH_eq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_eq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xH_eq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_H_vpeq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_vpeq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xpeq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xpeq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_peq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_peq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_peq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_vmeq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_vmeq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xmeq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xmeq_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_meq_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_meq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_meq_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_veqm_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_veqm_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xeqm_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xeqm_c_times_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xH_eqm_c_times_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_eqm_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_eqm_c_times_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_veq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_veq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_D_xeq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_D_xeq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xD_eq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_eq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xD_eq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_D_vpeq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_vpeq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xpeq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xpeq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_peq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_peq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_peq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_vmeq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_vmeq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xmeq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xmeq_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_meq_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_meq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_meq_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_veqm_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_veqm_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xeqm_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xeqm_c_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xD_eqm_c_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_eqm_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_eqm_c_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_veq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_M_xeq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_M_xeq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xM_eq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_eq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xM_eq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_M_vpeq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_vpeq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xpeq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xpeq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_peq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_peq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_peq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_vmeq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_vmeq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xmeq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xmeq_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_meq_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_meq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_meq_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veqm_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_veqm_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xeqm_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xeqm_c_times_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_eqm_c_times_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_eqm_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eqm_c_times_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_veq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xeq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xeq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xM_eq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_eq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_vpeq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_vpeq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xpeq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xpeq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_peq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_peq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_peq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_M_vmeq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_vmeq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xmeq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xmeq_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_meq_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_meq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_meq_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_M_veqm_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_veqm_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_xeqm_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xeqm_c_times_pMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xM_eqm_c_times_Ma(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_eqm_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xM_eqm_c_times_xMa(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_veq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_veq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_P_xeq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_P_xeq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_xP_eq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_eq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D_Complex *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xP_eq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_P_vpeq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_vpeq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xpeq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xpeq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_peq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_peq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_peq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_vmeq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_vmeq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xmeq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xmeq_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_meq_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_meq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_meq_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_veqm_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_veqm_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xeqm_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xeqm_c_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_eqm_c_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_eqm_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eqm_c_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_veq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_veq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xeq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xeq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_xP_eq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_eq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D_Complex *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_vpeq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_vpeq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xpeq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xpeq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_peq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_peq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_peq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_vmeq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_vmeq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xmeq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xmeq_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_meq_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_meq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_meq_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_P_veqm_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_veqm_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_P_xeqm_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_P_xeqm_c_times_pPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_xP_eqm_c_times_Pa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_P_eqm_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D_Complex *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_xP_eqm_c_times_xPa(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D_Complex *QLA_MI355X_RESTRICT a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);

/* D3: sum, r = a + b.
 * Reference: build_qla.pl:1421-1456, QLA_c_eq_c_plus_c. */
void QLA_D3_V_veq_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_V_veq_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_V_veq_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_veq_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_V_xeq_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_V_xeq_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_V_xeq_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_V_xeq_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_V_eq_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_V_eq_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_V_eq_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xV_eq_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_xV_eq_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xV_eq_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xV_eq_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_H_veq_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_H_veq_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_H_veq_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_veq_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_H_xeq_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_H_xeq_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_H_xeq_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_H_xeq_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_H_eq_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_H_eq_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_H_eq_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xH_eq_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_xH_eq_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xH_eq_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xH_eq_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_D_veq_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_D_veq_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_D_veq_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_veq_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_D_xeq_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_D_xeq_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_D_xeq_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_D_xeq_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_D_eq_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_D_eq_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_D_eq_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xD_eq_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_xD_eq_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xD_eq_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xD_eq_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_M_veq_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_M_veq_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_M_veq_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_veq_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_M_xeq_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_M_xeq_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_M_xeq_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_M_xeq_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_M_eq_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_M_eq_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_M_eq_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xM_eq_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_xM_eq_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xM_eq_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xM_eq_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_P_veq_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_P_veq_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_P_veq_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_veq_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_P_xeq_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_P_xeq_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_P_xeq_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_P_xeq_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_P_eq_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_P_eq_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_P_eq_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xP_eq_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_xP_eq_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_xP_eq_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xP_eq_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);

/* D3: difference, r = a - b.
 * Reference: build_qla.pl:1421-1456, QLA_c_eq_c_minus_c. */
void QLA_D3_V_veq_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_veq_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_V_veq_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_V_veq_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_V_xeq_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_V_xeq_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_V_xeq_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_V_xeq_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_V_eq_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_V_eq_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_V_eq_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_eq_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xV_eq_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xV_eq_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_xV_eq_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_H_veq_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_veq_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_H_veq_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_H_veq_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_H_xeq_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_H_xeq_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_H_xeq_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_H_xeq_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_H_eq_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_H_eq_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_H_eq_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_eq_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xH_eq_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xH_eq_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_xH_eq_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_D_veq_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_veq_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_D_veq_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_D_veq_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_D_xeq_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_D_xeq_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_D_xeq_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_D_xeq_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_D_eq_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_D_eq_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_D_eq_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_eq_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xD_eq_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xD_eq_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_xD_eq_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_veq_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_veq_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_M_veq_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_M_veq_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_M_xeq_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_M_xeq_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_M_xeq_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_M_xeq_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_M_eq_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_M_eq_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_M_eq_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eq_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xM_eq_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xM_eq_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_xM_eq_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_P_veq_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_veq_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_D3_P_veq_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_P_veq_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_P_xeq_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_D3_P_xeq_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_D3_P_xeq_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_P_xeq_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_P_eq_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_D3_P_eq_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_P_eq_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eq_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_D3_xP_eq_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_xP_eq_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_xP_eq_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);

/* D3: r = a[0] b + c, real a.
 * Reference: build_qla.pl:1837-1887, QLA_c_eq_r_times_c_plus_c. */
void QLA_D3_V_veq_r_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_V_veq_r_times_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_V_veq_r_times_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_veq_r_times_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_V_xeq_r_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_V_xeq_r_times_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_r_times_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_r_times_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_V_eq_r_times_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_V_eq_r_times_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_V_eq_r_times_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xV_eq_r_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xV_eq_r_times_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xV_eq_r_times_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xV_eq_r_times_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_H_veq_r_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_H_veq_r_times_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_H_veq_r_times_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_veq_r_times_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_H_xeq_r_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_H_xeq_r_times_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_r_times_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_r_times_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_H_eq_r_times_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_H_eq_r_times_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_H_eq_r_times_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xH_eq_r_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xH_eq_r_times_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xH_eq_r_times_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xH_eq_r_times_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_D_veq_r_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_D_veq_r_times_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_D_veq_r_times_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_veq_r_times_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_D_xeq_r_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_D_xeq_r_times_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_r_times_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_r_times_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_D_eq_r_times_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_D_eq_r_times_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_D_eq_r_times_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xD_eq_r_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xD_eq_r_times_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xD_eq_r_times_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xD_eq_r_times_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_M_veq_r_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_M_veq_r_times_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_M_veq_r_times_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_veq_r_times_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_M_xeq_r_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_M_xeq_r_times_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_r_times_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_r_times_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_M_eq_r_times_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_M_eq_r_times_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_M_eq_r_times_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xM_eq_r_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xM_eq_r_times_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xM_eq_r_times_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xM_eq_r_times_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_P_veq_r_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_P_veq_r_times_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_P_veq_r_times_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_veq_r_times_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_P_xeq_r_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_P_xeq_r_times_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_r_times_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_r_times_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_P_eq_r_times_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_P_eq_r_times_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_P_eq_r_times_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xP_eq_r_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Real *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xP_eq_r_times_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xP_eq_r_times_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xP_eq_r_times_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);

/* D3: r = a[0] b - c, real a.
 * Reference: build_qla.pl:1837-1887, QLA_c_eq_r_times_c_minus_c. */
void QLA_D3_V_veq_r_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_veq_r_times_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_V_veq_r_times_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_V_veq_r_times_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_V_xeq_r_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_r_times_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_V_xeq_r_times_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_V_xeq_r_times_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_V_eq_r_times_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_V_eq_r_times_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_eq_r_times_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xV_eq_r_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xV_eq_r_times_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xV_eq_r_times_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xV_eq_r_times_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_H_veq_r_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_veq_r_times_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_H_veq_r_times_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_H_veq_r_times_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_H_xeq_r_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_r_times_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_H_xeq_r_times_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_H_xeq_r_times_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_H_eq_r_times_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_H_eq_r_times_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_eq_r_times_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xH_eq_r_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xH_eq_r_times_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xH_eq_r_times_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xH_eq_r_times_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_D_veq_r_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_veq_r_times_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_D_veq_r_times_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_D_veq_r_times_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_D_xeq_r_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_r_times_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_D_xeq_r_times_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_D_xeq_r_times_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_D_eq_r_times_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_D_eq_r_times_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_eq_r_times_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xD_eq_r_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xD_eq_r_times_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xD_eq_r_times_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xD_eq_r_times_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_M_veq_r_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_veq_r_times_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_M_veq_r_times_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_M_veq_r_times_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_M_xeq_r_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_r_times_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_M_xeq_r_times_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_M_xeq_r_times_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_M_eq_r_times_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_M_eq_r_times_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_eq_r_times_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xM_eq_r_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xM_eq_r_times_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xM_eq_r_times_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xM_eq_r_times_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_P_veq_r_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_veq_r_times_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_P_veq_r_times_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_P_veq_r_times_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_P_xeq_r_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_r_times_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_P_xeq_r_times_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_P_xeq_r_times_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_P_eq_r_times_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_P_eq_r_times_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_eq_r_times_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xP_eq_r_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Real *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xP_eq_r_times_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xP_eq_r_times_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Real *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xP_eq_r_times_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Real *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);

/* D3: r = a[0] b + c, complex a.
 * Reference: build_qla.pl:1889-1935, QLA_c_eq_c_times_c_plus_c. */
void QLA_D3_V_veq_c_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_V_veq_c_times_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_V_veq_c_times_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_veq_c_times_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_V_xeq_c_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_V_xeq_c_times_V_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_c_times_pV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_c_times_pV_plus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_V_eq_c_times_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_V_eq_c_times_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_V_eq_c_times_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xV_eq_c_times_V_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xV_eq_c_times_V_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xV_eq_c_times_xV_plus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xV_eq_c_times_xV_plus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_H_veq_c_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_H_veq_c_times_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_H_veq_c_times_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_veq_c_times_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_H_xeq_c_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_H_xeq_c_times_H_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_c_times_pH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_c_times_pH_plus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_H_eq_c_times_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_H_eq_c_times_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_H_eq_c_times_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xH_eq_c_times_H_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xH_eq_c_times_H_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xH_eq_c_times_xH_plus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xH_eq_c_times_xH_plus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_D_veq_c_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_D_veq_c_times_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_D_veq_c_times_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_veq_c_times_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_D_xeq_c_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_D_xeq_c_times_D_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_c_times_pD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_c_times_pD_plus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_D_eq_c_times_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_D_eq_c_times_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_D_eq_c_times_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xD_eq_c_times_D_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xD_eq_c_times_D_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xD_eq_c_times_xD_plus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xD_eq_c_times_xD_plus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_M_veq_c_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_M_veq_c_times_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_M_veq_c_times_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_veq_c_times_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_M_xeq_c_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_M_xeq_c_times_M_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_c_times_pM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_c_times_pM_plus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_M_eq_c_times_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_M_eq_c_times_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_M_eq_c_times_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xM_eq_c_times_M_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                   QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xM_eq_c_times_M_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xM_eq_c_times_xM_plus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xM_eq_c_times_xM_plus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_P_veq_c_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_P_veq_c_times_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int n);
void QLA_D3_P_veq_c_times_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_veq_c_times_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_P_xeq_c_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *index,
                                   int n);
void QLA_D3_P_xeq_c_times_P_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_c_times_pP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_c_times_pP_plus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_P_eq_c_times_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int *indexc,
                                   int n);
void QLA_D3_P_eq_c_times_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   int *indexb,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_P_eq_c_times_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xP_eq_c_times_P_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                   int *indexr,
                                   QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                   QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                   int n);
void QLA_D3_xP_eq_c_times_P_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_xP_eq_c_times_xP_plus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xP_eq_c_times_xP_plus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);

/* D3: r = a[0] b - c, complex a.
 * Reference: build_qla.pl:1889-1935, QLA_c_eq_c_times_c_minus_c. */
void QLA_D3_V_veq_c_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_veq_c_times_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_V_veq_c_times_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_V_veq_c_times_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_V_xeq_c_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_V_xeq_c_times_V_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_V_xeq_c_times_pV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_V_xeq_c_times_pV_minus_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_V_eq_c_times_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_V_eq_c_times_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_V_eq_c_times_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xV_eq_c_times_V_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xV_eq_c_times_V_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xV_eq_c_times_xV_minus_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xV_eq_c_times_xV_minus_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_H_veq_c_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_veq_c_times_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_H_veq_c_times_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_H_veq_c_times_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_H_xeq_c_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_H_xeq_c_times_H_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_H_xeq_c_times_pH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_H_xeq_c_times_pH_minus_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_H_eq_c_times_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_H_eq_c_times_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_H_eq_c_times_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xH_eq_c_times_H_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xH_eq_c_times_H_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xH_eq_c_times_xH_minus_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xH_eq_c_times_xH_minus_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_D_veq_c_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_veq_c_times_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_D_veq_c_times_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_D_veq_c_times_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_D_xeq_c_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_D_xeq_c_times_D_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_D_xeq_c_times_pD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_D_xeq_c_times_pD_minus_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_D_eq_c_times_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_D_eq_c_times_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_D_eq_c_times_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xD_eq_c_times_D_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xD_eq_c_times_D_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xD_eq_c_times_xD_minus_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xD_eq_c_times_xD_minus_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_M_veq_c_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_veq_c_times_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_M_veq_c_times_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_M_veq_c_times_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_M_xeq_c_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_M_xeq_c_times_M_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_M_xeq_c_times_pM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_M_xeq_c_times_pM_minus_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_M_eq_c_times_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_M_eq_c_times_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_M_eq_c_times_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xM_eq_c_times_M_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xM_eq_c_times_M_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xM_eq_c_times_xM_minus_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xM_eq_c_times_xM_minus_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);
void QLA_D3_P_veq_c_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_veq_c_times_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int n);
void QLA_D3_P_veq_c_times_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_P_veq_c_times_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int n);
void QLA_D3_P_xeq_c_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *index,
                                    int n);
void QLA_D3_P_xeq_c_times_P_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                     int *index,
                                     int n);
void QLA_D3_P_xeq_c_times_pP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *index,
                                     int n);
void QLA_D3_P_xeq_c_times_pP_minus_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *c,
                                      int *index,
                                      int n);
void QLA_D3_P_eq_c_times_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int *indexc,
                                    int n);
void QLA_D3_P_eq_c_times_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    int *indexb,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_P_eq_c_times_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xP_eq_c_times_P_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                    int *indexr,
                                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                    int n);
void QLA_D3_xP_eq_c_times_P_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int *indexc,
                                     int n);
void QLA_D3_xP_eq_c_times_xP_minus_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                     int *indexr,
                                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                     int *indexb,
                                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                     int n);
void QLA_D3_xP_eq_c_times_xP_minus_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                      int *indexr,
                                      QLA_D_Complex *QLA_MI355X_RESTRICT a,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                                      int *indexb,
                                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT c,
                                      int *indexc,
                                      int n);

/* F3: zero fill, r = 0.
 * Reference: build_qla.pl:2561-2580, QLA_c_eq_r(e, 0.). */
void QLA_F3_V_veq_zero(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_F3_V_xeq_zero(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_F3_H_veq_zero(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_F3_H_xeq_zero(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_F3_D_veq_zero(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_F3_D_xeq_zero(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_F3_M_veq_zero(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_F3_M_xeq_zero(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_F3_P_veq_zero(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_F3_P_xeq_zero(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);

/* F3: constant fill, r eqop a[0] (one record for every site).
 * Reference: build_qla.pl:2584-2603, QLA_c_<eqop>_c. */
void QLA_F3_V_veq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_V_xeq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_V_vpeq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_xpeq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_V_vmeq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_xmeq_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_V_veqm_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_V_xeqm_v(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_veq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_H_xeq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_H_vpeq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_xpeq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_vmeq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_xmeq_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_H_veqm_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_H_xeqm_h(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_veq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_D_xeq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_D_vpeq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_xpeq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_vmeq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_xmeq_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_D_veqm_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_D_xeqm_d(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_veq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_M_xeq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_M_vpeq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xpeq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_vmeq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xmeq_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_veqm_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xeqm_m(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_veq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_P_xeq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_P_vpeq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_xpeq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_vmeq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_xmeq_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_P_veqm_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_P_xeqm_p(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);

/* F3: diagonal fill, r eqop a[0] times the unit matrix.
 * Reference: build_qla.pl:2605-2617, expressions_tensor.pl:77-128 (QLA_c_<eqop>_r(e, 0.), then the diagonal). */
void QLA_F3_M_veq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_F3_M_xeq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_F_Complex *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_F3_M_vpeq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xpeq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_vmeq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xmeq_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_F3_M_veqm_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_F3_M_xeqm_c(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);

/* D3: zero fill, r = 0.
 * Reference: build_qla.pl:2561-2580, QLA_c_eq_r(e, 0.). */
void QLA_D3_V_veq_zero(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_D3_V_xeq_zero(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_D3_H_veq_zero(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_D3_H_xeq_zero(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_D3_D_veq_zero(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_D3_D_xeq_zero(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_D3_M_veq_zero(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_D3_M_xeq_zero(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);
void QLA_D3_P_veq_zero(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int n);
void QLA_D3_P_xeq_zero(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                       int *index,
                       int n);

/* D3: constant fill, r eqop a[0] (one record for every site).
 * Reference: build_qla.pl:2584-2603, QLA_c_<eqop>_c. */
void QLA_D3_V_veq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_V_xeq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_V_vpeq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_xpeq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_V_vmeq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_xmeq_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_V_veqm_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_V_xeqm_v(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_veq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_H_xeq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_H_vpeq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_xpeq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_vmeq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_xmeq_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_H_veqm_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_H_xeqm_h(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_veq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_D_xeq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_D_vpeq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_xpeq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_vmeq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_xmeq_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_D_veqm_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_D_xeqm_d(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_veq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_M_xeq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_M_vpeq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xpeq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_vmeq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xmeq_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_veqm_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xeqm_m(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_veq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_P_xeq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                    QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_P_vpeq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_xpeq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_vmeq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_xmeq_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_P_veqm_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_P_xeqm_p(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);

/* D3: diagonal fill, r eqop a[0] times the unit matrix.
 * Reference: build_qla.pl:2605-2617, expressions_tensor.pl:77-128 (QLA_c_<eqop>_r(e, 0.), then the diagonal). */
void QLA_D3_M_veq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                    int n);
void QLA_D3_M_xeq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                    QLA_D_Complex *QLA_MI355X_RESTRICT a,
                    int *index,
                    int n);
void QLA_D3_M_vpeq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xpeq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_vmeq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xmeq_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_D3_M_veqm_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_D3_M_xeqm_c(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D_Complex *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);

/* F3: left multiplication by Gamma_mu = g0^b0 g1^b1 g2^b2 g3^b3 (bits b of mu = 0..15), r = Gamma_mu a.
 * Reference: build_qla.pl:1356-1384, expressions_gamma.pl:123-198, 387-470. */
void QLA_F3_D_veq_gamma_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_D_veq_gamma_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_F3_D_xeq_gamma_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_F3_D_xeq_gamma_times_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_F3_xD_eq_gamma_times_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_D_eq_gamma_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_F3_xD_eq_gamma_times_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);
void QLA_F3_P_veq_gamma_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_P_veq_gamma_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_F3_P_xeq_gamma_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_F3_P_xeq_gamma_times_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_F3_xP_eq_gamma_times_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_P_eq_gamma_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_F3_xP_eq_gamma_times_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);

/* F3: right multiplication by Gamma_mu, r = a Gamma_mu.
 * Reference: build_qla.pl:1386-1415, expressions_gamma.pl. */
void QLA_F3_P_veq_P_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_P_veq_pP_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_F3_P_xeq_P_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_F3_P_xeq_pP_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_F3_xP_eq_P_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_F3_P_eq_xP_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_F3_xP_eq_xP_times_gamma(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);

/* D3: left multiplication by Gamma_mu = g0^b0 g1^b1 g2^b2 g3^b3 (bits b of mu = 0..15), r = Gamma_mu a.
 * Reference: build_qla.pl:1356-1384, expressions_gamma.pl:123-198, 387-470. */
void QLA_D3_D_veq_gamma_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_D_veq_gamma_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_D3_D_xeq_gamma_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_D3_D_xeq_gamma_times_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_D3_xD_eq_gamma_times_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_D_eq_gamma_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_D3_xD_eq_gamma_times_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);
void QLA_D3_P_veq_gamma_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_P_veq_gamma_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_D3_P_xeq_gamma_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_D3_P_xeq_gamma_times_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_D3_xP_eq_gamma_times_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_P_eq_gamma_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_D3_xP_eq_gamma_times_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);

/* D3: right multiplication by Gamma_mu, r = a Gamma_mu.
 * Reference: build_qla.pl:1386-1415, expressions_gamma.pl. */
void QLA_D3_P_veq_P_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_P_veq_pP_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int n);
void QLA_D3_P_xeq_P_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int *index,
                                int n);
void QLA_D3_P_xeq_pP_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                 int mu,
                                 int *index,
                                 int n);
void QLA_D3_xP_eq_P_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int mu,
                                int n);
void QLA_D3_P_eq_xP_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int mu,
                                int n);
void QLA_D3_xP_eq_xP_times_gamma(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                                 int *indexr,
                                 QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                                 int *indexa,
                                 int mu,
                                 int n);

#endif
