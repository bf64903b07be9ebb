/*
 * qla_mi355x_fields.h -- global reductions and precision conversions of the QLA
 * C ABI on MI355X (types V, H, D, M, P).
 * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.
 * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device
 * semantics are documented there).
 *
 * Names and argument order are QLA's (lib/perl/prototype.pl:450-688).  A `p`
 * operand is a pointer array; x-forms use the records at index[i].  The result
 * pointer of a reduction may be a host or a device address; n <= 0 gives 0.
 * QLA_DF3_* reductions take float fields and return double results; QLA_DF3_*
 * / QLA_FD3_* conversions write double / float fields.
 */
#ifndef QLA_MI355X_FIELDS_H
#define QLA_MI355X_FIELDS_H

/* F3: global squared norm, sum over sites and components of |a|^2.
 * Reference: build_qla.pl:2089-2156, make_code_unary.pl:318-366; per component the norm is formed in the field precision and added to a double sum. */
void QLA_F3_r_veq_norm2_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_r_veq_norm2_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_r_xeq_norm2_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_r_xeq_norm2_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_r_veq_norm2_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_r_veq_norm2_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_r_xeq_norm2_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_r_xeq_norm2_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_r_veq_norm2_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_r_veq_norm2_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_r_xeq_norm2_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_r_xeq_norm2_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_r_veq_norm2_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_r_veq_norm2_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_r_xeq_norm2_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_r_xeq_norm2_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_r_veq_norm2_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_r_veq_norm2_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_r_xeq_norm2_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_r_xeq_norm2_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);

/* F3: global inner product, sum of conj(a) b over sites and components.
 * Reference: build_qla.pl:2196-2360, make_code_binary.pl:110-170; operands promoted to double. */
void QLA_F3_c_veq_V_dot_V(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_veq_V_dot_pV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_F3_c_veq_pV_dot_V(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_c_veq_pV_dot_pV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_c_xeq_V_dot_V(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_F3_c_xeq_V_dot_pV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pV_dot_V(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pV_dot_pV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_c_eq_xV_dot_V(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_eq_V_dot_xV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_F3_c_eq_xV_dot_xV(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_c_veq_H_dot_H(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_veq_H_dot_pH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_F3_c_veq_pH_dot_H(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_c_veq_pH_dot_pH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_c_xeq_H_dot_H(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_F3_c_xeq_H_dot_pH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pH_dot_H(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pH_dot_pH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_c_eq_xH_dot_H(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_eq_H_dot_xH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_F3_c_eq_xH_dot_xH(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_c_veq_D_dot_D(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_veq_D_dot_pD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_F3_c_veq_pD_dot_D(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_c_veq_pD_dot_pD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_c_xeq_D_dot_D(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_F3_c_xeq_D_dot_pD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pD_dot_D(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pD_dot_pD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_c_eq_xD_dot_D(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_eq_D_dot_xD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_F3_c_eq_xD_dot_xD(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_c_veq_M_dot_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_veq_M_dot_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_F3_c_veq_pM_dot_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_c_veq_pM_dot_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_c_xeq_M_dot_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_F3_c_xeq_M_dot_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pM_dot_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pM_dot_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_c_eq_xM_dot_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_eq_M_dot_xM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_F3_c_eq_xM_dot_xM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_F3_c_veq_P_dot_P(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_veq_P_dot_pP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_F3_c_veq_pP_dot_P(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_F3_c_veq_pP_dot_pP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_F3_c_xeq_P_dot_P(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_F3_c_xeq_P_dot_pP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pP_dot_P(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_F3_c_xeq_pP_dot_pP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_F3_c_eq_xP_dot_P(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_F3_c_eq_P_dot_xP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_F3_c_eq_xP_dot_xP(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);

/* F3: real part of the global inner product.
 * Reference: build_qla.pl:2244-2260 (qualifier re), make_code_binary.pl:110-170. */
void QLA_F3_r_veq_re_V_dot_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_veq_re_V_dot_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_r_veq_re_pV_dot_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_r_veq_re_pV_dot_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_r_xeq_re_V_dot_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_r_xeq_re_V_dot_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pV_dot_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pV_dot_pV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_r_eq_re_xV_dot_V(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_eq_re_V_dot_xV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_r_eq_re_xV_dot_xV(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_r_veq_re_H_dot_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_veq_re_H_dot_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_r_veq_re_pH_dot_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_r_veq_re_pH_dot_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_r_xeq_re_H_dot_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_r_xeq_re_H_dot_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pH_dot_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pH_dot_pH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_r_eq_re_xH_dot_H(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_eq_re_H_dot_xH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_r_eq_re_xH_dot_xH(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_r_veq_re_D_dot_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_veq_re_D_dot_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_r_veq_re_pD_dot_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_r_veq_re_pD_dot_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_r_xeq_re_D_dot_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_r_xeq_re_D_dot_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pD_dot_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pD_dot_pD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_r_eq_re_xD_dot_D(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_eq_re_D_dot_xD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_r_eq_re_xD_dot_xD(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_r_veq_re_M_dot_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_veq_re_M_dot_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_r_veq_re_pM_dot_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_r_veq_re_pM_dot_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_r_xeq_re_M_dot_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_r_xeq_re_M_dot_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pM_dot_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pM_dot_pM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_r_eq_re_xM_dot_M(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_eq_re_M_dot_xM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_r_eq_re_xM_dot_xM(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_r_veq_re_P_dot_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_veq_re_P_dot_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_F3_r_veq_re_pP_dot_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_r_veq_re_pP_dot_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_r_xeq_re_P_dot_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_F3_r_xeq_re_P_dot_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pP_dot_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_r_xeq_re_pP_dot_pP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_r_eq_re_xP_dot_P(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_F3_r_eq_re_P_dot_xP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_F3_r_eq_re_xP_dot_xP(QLA_F_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);

/* D3: global squared norm, sum over sites and components of |a|^2.
 * Reference: build_qla.pl:2089-2156, make_code_unary.pl:318-366; per component the norm is formed in the field precision and added to a double sum. */
void QLA_D3_r_veq_norm2_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_r_veq_norm2_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_r_xeq_norm2_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_r_xeq_norm2_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_r_veq_norm2_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_r_veq_norm2_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_r_xeq_norm2_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_r_xeq_norm2_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_r_veq_norm2_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_r_veq_norm2_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_r_xeq_norm2_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_r_xeq_norm2_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_r_veq_norm2_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_r_veq_norm2_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_r_xeq_norm2_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_r_xeq_norm2_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_r_veq_norm2_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_r_veq_norm2_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_r_xeq_norm2_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_r_xeq_norm2_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);

/* D3: global inner product, sum of conj(a) b over sites and components.
 * Reference: build_qla.pl:2196-2360, make_code_binary.pl:110-170; operands promoted to double. */
void QLA_D3_c_veq_V_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_veq_V_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_D3_c_veq_pV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_c_veq_pV_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_c_xeq_V_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_D3_c_xeq_V_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pV_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_c_eq_xV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_eq_V_dot_xV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_D3_c_eq_xV_dot_xV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_c_veq_H_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_veq_H_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_D3_c_veq_pH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_c_veq_pH_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_c_xeq_H_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_D3_c_xeq_H_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pH_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_c_eq_xH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_eq_H_dot_xH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_D3_c_eq_xH_dot_xH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_c_veq_D_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_veq_D_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_D3_c_veq_pD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_c_veq_pD_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_c_xeq_D_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_D3_c_xeq_D_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pD_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_c_eq_xD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_eq_D_dot_xD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_D3_c_eq_xD_dot_xD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_c_veq_M_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_veq_M_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_D3_c_veq_pM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_c_veq_pM_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_c_xeq_M_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_D3_c_xeq_M_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pM_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_c_eq_xM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_eq_M_dot_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_D3_c_eq_xM_dot_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_D3_c_veq_P_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_veq_P_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                           int n);
void QLA_D3_c_veq_pP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_D3_c_veq_pP_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_D3_c_xeq_P_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int *index,
                          int n);
void QLA_D3_c_xeq_P_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_D3_c_xeq_pP_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_D3_c_eq_xP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int n);
void QLA_D3_c_eq_P_dot_xP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                          QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                          int *indexb,
                          int n);
void QLA_D3_c_eq_xP_dot_xP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);

/* D3: real part of the global inner product.
 * Reference: build_qla.pl:2244-2260 (qualifier re), make_code_binary.pl:110-170. */
void QLA_D3_r_veq_re_V_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_veq_re_V_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_r_veq_re_pV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_r_veq_re_pV_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_r_xeq_re_V_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_r_xeq_re_V_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pV_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_r_eq_re_xV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_eq_re_V_dot_xV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_r_eq_re_xV_dot_xV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_r_veq_re_H_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_veq_re_H_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_r_veq_re_pH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_r_veq_re_pH_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_r_xeq_re_H_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_r_xeq_re_H_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pH_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_r_eq_re_xH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_eq_re_H_dot_xH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_r_eq_re_xH_dot_xH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_r_veq_re_D_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_veq_re_D_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_r_veq_re_pD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_r_veq_re_pD_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_r_xeq_re_D_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_r_xeq_re_D_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pD_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_r_eq_re_xD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_eq_re_D_dot_xD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_r_eq_re_xD_dot_xD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_r_veq_re_M_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_veq_re_M_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_r_veq_re_pM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_r_veq_re_pM_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_r_xeq_re_M_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_r_xeq_re_M_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pM_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_r_eq_re_xM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_eq_re_M_dot_xM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_r_eq_re_xM_dot_xM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_r_veq_re_P_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_veq_re_P_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int n);
void QLA_D3_r_veq_re_pP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_r_veq_re_pP_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_r_xeq_re_P_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *index,
                             int n);
void QLA_D3_r_xeq_re_P_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_r_xeq_re_pP_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_r_eq_re_xP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int n);
void QLA_D3_r_eq_re_P_dot_xP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                             QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                             int *indexb,
                             int n);
void QLA_D3_r_eq_re_xP_dot_xP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);

/* DF3: global squared norm, sum over sites and components of |a|^2.
 * Reference: build_qla.pl:2089-2156, make_code_unary.pl:318-366; per component the norm is formed in the field precision and added to a double sum. */
void QLA_DF3_r_veq_norm2_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_DF3_r_veq_norm2_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_DF3_r_xeq_norm2_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_DF3_r_xeq_norm2_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_DF3_r_veq_norm2_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_DF3_r_veq_norm2_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_DF3_r_xeq_norm2_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_DF3_r_xeq_norm2_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_DF3_r_veq_norm2_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_DF3_r_veq_norm2_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_DF3_r_xeq_norm2_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_DF3_r_xeq_norm2_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_DF3_r_veq_norm2_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_DF3_r_veq_norm2_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_DF3_r_xeq_norm2_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_DF3_r_xeq_norm2_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_DF3_r_veq_norm2_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_DF3_r_veq_norm2_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_DF3_r_xeq_norm2_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_DF3_r_xeq_norm2_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);

/* DF3: global inner product, sum of conj(a) b over sites and components.
 * Reference: build_qla.pl:2196-2360, make_code_binary.pl:110-170; operands promoted to double. */
void QLA_DF3_c_veq_V_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_veq_V_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_DF3_c_veq_pV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_DF3_c_veq_pV_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_DF3_c_xeq_V_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_DF3_c_xeq_V_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pV_dot_pV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_DF3_c_eq_xV_dot_V(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_eq_V_dot_xV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_DF3_c_eq_xV_dot_xV(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_DF3_c_veq_H_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_veq_H_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_DF3_c_veq_pH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_DF3_c_veq_pH_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_DF3_c_xeq_H_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_DF3_c_xeq_H_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pH_dot_pH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_DF3_c_eq_xH_dot_H(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_eq_H_dot_xH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_DF3_c_eq_xH_dot_xH(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_DF3_c_veq_D_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_veq_D_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_DF3_c_veq_pD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_DF3_c_veq_pD_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_DF3_c_xeq_D_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_DF3_c_xeq_D_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pD_dot_pD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_DF3_c_eq_xD_dot_D(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_eq_D_dot_xD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_DF3_c_eq_xD_dot_xD(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_DF3_c_veq_M_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_veq_M_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_DF3_c_veq_pM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_DF3_c_veq_pM_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_DF3_c_xeq_M_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_DF3_c_xeq_M_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pM_dot_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_DF3_c_eq_xM_dot_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_eq_M_dot_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_DF3_c_eq_xM_dot_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);
void QLA_DF3_c_veq_P_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_veq_P_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int n);
void QLA_DF3_c_veq_pP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int n);
void QLA_DF3_c_veq_pP_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int n);
void QLA_DF3_c_xeq_P_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *index,
                           int n);
void QLA_DF3_c_xeq_P_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *index,
                            int n);
void QLA_DF3_c_xeq_pP_dot_pP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                             QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                             int *index,
                             int n);
void QLA_DF3_c_eq_xP_dot_P(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int n);
void QLA_DF3_c_eq_P_dot_xP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                           QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                           int *indexb,
                           int n);
void QLA_DF3_c_eq_xP_dot_xP(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                            int *indexb,
                            int n);

/* DF3: real part of the global inner product.
 * Reference: build_qla.pl:2244-2260 (qualifier re), make_code_binary.pl:110-170. */
void QLA_DF3_r_veq_re_V_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_veq_re_V_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_DF3_r_veq_re_pV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_DF3_r_veq_re_pV_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_DF3_r_xeq_re_V_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_DF3_r_xeq_re_V_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pV_dot_pV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_DF3_r_eq_re_xV_dot_V(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_eq_re_V_dot_xV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_DF3_r_eq_re_xV_dot_xV(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_DF3_r_veq_re_H_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_veq_re_H_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_DF3_r_veq_re_pH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_DF3_r_veq_re_pH_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                                QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_DF3_r_xeq_re_H_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_DF3_r_xeq_re_H_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pH_dot_pH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                                QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_DF3_r_eq_re_xH_dot_H(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_eq_re_H_dot_xH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_DF3_r_eq_re_xH_dot_xH(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_HalfFermion *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_DF3_r_veq_re_D_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_veq_re_D_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_DF3_r_veq_re_pD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_DF3_r_veq_re_pD_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_DF3_r_xeq_re_D_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_DF3_r_xeq_re_D_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pD_dot_pD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                                QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_DF3_r_eq_re_xD_dot_D(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_eq_re_D_dot_xD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_DF3_r_eq_re_xD_dot_xD(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_DiracFermion *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_DF3_r_veq_re_M_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_veq_re_M_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_DF3_r_veq_re_pM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_DF3_r_veq_re_pM_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_DF3_r_xeq_re_M_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_DF3_r_xeq_re_M_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pM_dot_pM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_DF3_r_eq_re_xM_dot_M(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_eq_re_M_dot_xM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_DF3_r_eq_re_xM_dot_xM(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_DF3_r_veq_re_P_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_veq_re_P_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_DF3_r_veq_re_pP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_DF3_r_veq_re_pP_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_DF3_r_xeq_re_P_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_DF3_r_xeq_re_P_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_DF3_r_xeq_re_pP_dot_pP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                                QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_DF3_r_eq_re_xP_dot_P(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_DF3_r_eq_re_P_dot_xP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                              QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_DF3_r_eq_re_xP_dot_xP(QLA_D_Real *QLA_MI355X_RESTRICT r,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);

/* DF3: precision conversion of a field (DF3: float to double, exact; FD3: double to float, rounded to nearest).
 * Reference: build_qla.pl:530-585, make_code_unary.pl (identity with a cast). */
void QLA_DF3_V_veq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_V_veq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_DF3_V_xeq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_DF3_V_xeq_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_DF3_xV_eq_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_V_eq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_DF3_xV_eq_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_DF3_H_veq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_H_veq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_DF3_H_xeq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_DF3_H_xeq_pH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_DF3_xH_eq_H(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_H_eq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_DF3_xH_eq_xH(QLA_D3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_DF3_D_veq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_D_veq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_DF3_D_xeq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_DF3_D_xeq_pD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_DF3_xD_eq_D(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_D_eq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_DF3_xD_eq_xD(QLA_D3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_DF3_M_veq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_M_veq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_DF3_M_xeq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_DF3_M_xeq_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_DF3_xM_eq_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_M_eq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_DF3_xM_eq_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_DF3_P_veq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_P_veq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_DF3_P_xeq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_DF3_P_xeq_pP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_DF3_xP_eq_P(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_DF3_P_eq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_DF3_xP_eq_xP(QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);

/* FD3: precision conversion of a field (DF3: float to double, exact; FD3: double to float, rounded to nearest).
 * Reference: build_qla.pl:530-585, make_code_unary.pl (identity with a cast). */
void QLA_FD3_V_veq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_V_veq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_FD3_V_xeq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_FD3_V_xeq_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_FD3_xV_eq_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_V_eq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_FD3_xV_eq_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorVector *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_FD3_H_veq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_H_veq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_FD3_H_xeq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_FD3_H_xeq_pH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_FD3_xH_eq_H(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_H_eq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_FD3_xH_eq_xH(QLA_F3_HalfFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_HalfFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_FD3_D_veq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_D_veq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_FD3_D_xeq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_FD3_D_xeq_pD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_FD3_xD_eq_D(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_D_eq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_FD3_xD_eq_xD(QLA_F3_DiracFermion *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracFermion *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_FD3_M_veq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_M_veq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_FD3_M_xeq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_FD3_M_xeq_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_FD3_xM_eq_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_M_eq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                     QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_FD3_xM_eq_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);
void QLA_FD3_P_veq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_P_veq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int n);
void QLA_FD3_P_xeq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *index,
                     int n);
void QLA_FD3_P_xeq_pP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT *a,
                      int *index,
                      int n);
void QLA_FD3_xP_eq_P(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     int *indexr,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int n);
void QLA_FD3_P_eq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                     QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                     int *indexa,
                     int n);
void QLA_FD3_xP_eq_xP(QLA_F3_DiracPropagator *QLA_MI355X_RESTRICT r,
                      int *indexr,
                      QLA_D3_DiracPropagator *QLA_MI355X_RESTRICT a,
                      int *indexa,
                      int n);

#endif
