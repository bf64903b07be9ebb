/*
 * qla_mi355x_matfuncs.h -- colour-matrix functions of the QLA C ABI on MI355X.
 * GENERATED by qla_amd/abi_gen.py from its variant table; do not edit.
 * Included by qla_mi355x.h (types, QLA_MI355X_RESTRICT and the device
 * semantics are documented there).
 *
 * Names and argument order are QLA's (lib/perl/build_qla.pl:1002-1066 and
 * 1663-1690, prototype.pl:450-688).  QLA_FN_* take the number of colours
 * first (nc = 3..8 here; other values record an error and write nothing)
 * and their record arguments are untyped: nc x nc complex floats per matrix,
 * nc per vector, or a pointer array of such records for a `p` operand.  A `p`
 * operand is a pointer array; x-forms use every operand at index[i].
 */
#ifndef QLA_MI355X_MATFUNCS_H
#define QLA_MI355X_MATFUNCS_H

/* F3: matrix exponential.
 * Reference: code-templates/M_eq_exp_M.c (scaled Pade(3,3), squaring). */
void QLA_F3_M_veq_exp_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_M_veq_exp_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_F3_M_xeq_exp_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_F3_M_xeq_exp_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_F3_xM_eq_exp_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_M_eq_exp_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_xM_eq_exp_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* F3: matrix square root.
 * Reference: code-templates/M_eq_sqrt_M.c (Newton iteration). */
void QLA_F3_M_veq_sqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_veq_sqrt_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_M_xeq_sqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_M_xeq_sqrt_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xM_eq_sqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_eq_sqrt_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xM_eq_sqrt_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);

/* F3: inverse square root.
 * Reference: code-templates/M_eq_invsqrt_M.c (Newton iteration). */
void QLA_F3_M_veq_invsqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_F3_M_veq_invsqrt_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int n);
void QLA_F3_M_xeq_invsqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *index,
                            int n);
void QLA_F3_M_xeq_invsqrt_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int *index,
                             int n);
void QLA_F3_xM_eq_invsqrt_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_F3_M_eq_invsqrt_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);
void QLA_F3_xM_eq_invsqrt_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             int n);

/* F3: determinant.
 * Reference: code-templates/C_eq_det_M.c (closed form for nc = 3, pivoted LU otherwise). */
void QLA_F3_C_veq_det_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_C_veq_det_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_F3_C_xeq_det_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_F3_C_xeq_det_pM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_F3_xC_eq_det_M(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_C_eq_det_xM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_xC_eq_det_xM(QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* F3: matrix inverse.
 * Reference: code-templates/M_eq_inverse_M.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_F3_M_veq_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_F3_M_veq_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int n);
void QLA_F3_M_xeq_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *index,
                            int n);
void QLA_F3_M_xeq_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int *index,
                             int n);
void QLA_F3_xM_eq_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_F3_M_eq_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);
void QLA_F3_xM_eq_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             int n);

/* F3: a^-1 b for a colour vector b.
 * Reference: code-templates/V_eq_M_inverse_V.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_F3_V_veq_M_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_V_veq_M_inverse_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_V_veq_pM_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_F3_V_veq_pM_inverse_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_F3_V_xeq_M_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_V_xeq_M_inverse_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_V_xeq_pM_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_F3_V_xeq_pM_inverse_pV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_F3_V_eq_M_inverse_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_V_eq_xM_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_V_eq_xM_inverse_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_xV_eq_M_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_xV_eq_M_inverse_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_xV_eq_xM_inverse_V(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_F3_xV_eq_xM_inverse_xV(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                QLA_F3_ColorVector *QLA_MI355X_RESTRICT b,
                                int *indexb,
                                int n);

/* F3: a^-1 b for a colour matrix b.
 * Reference: code-templates/M_eq_M_inverse_M.c (Gauss-Jordan). */
void QLA_F3_M_veq_M_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_veq_M_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_F3_M_veq_pM_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_F3_M_veq_pM_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_F3_M_xeq_M_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_F3_M_xeq_M_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_F3_M_xeq_pM_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_F3_M_xeq_pM_inverse_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_F3_M_eq_M_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_F3_M_eq_xM_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_M_eq_xM_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_xM_eq_M_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_F3_xM_eq_M_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_F3_xM_eq_xM_inverse_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_F3_xM_eq_xM_inverse_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                int *indexb,
                                int n);

/* F3: exponential of an anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_F3_M_veq_expA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_veq_expA_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_F3_M_xeq_expA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_F3_M_xeq_expA_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_F3_xM_eq_expA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_F3_M_eq_expA_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_F3_xM_eq_expA_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);

/* F3: exponential of a traceless anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expTA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_F3_M_veq_expTA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_veq_expTA_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_F3_M_xeq_expTA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_F3_M_xeq_expTA_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_F3_xM_eq_expTA_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_F3_M_eq_expTA_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_F3_xM_eq_expTA_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);

/* F3: square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_sqrtPH_M.c (eigenvalue closed form for nc = 3, the sqrt branch otherwise). */
void QLA_F3_M_veq_sqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_F3_M_veq_sqrtPH_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_F3_M_xeq_sqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_F3_M_xeq_sqrtPH_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_F3_xM_eq_sqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_F3_M_eq_sqrtPH_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_F3_xM_eq_sqrtPH_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);

/* F3: inverse square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_invsqrtPH_M.c (closed form for nc = 3, the invsqrt branch otherwise). */
void QLA_F3_M_veq_invsqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_M_veq_invsqrtPH_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_F3_M_xeq_invsqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_F3_M_xeq_invsqrtPH_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_F3_xM_eq_invsqrtPH_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_F3_M_eq_invsqrtPH_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_F3_xM_eq_invsqrtPH_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);

/* F3: matrix logarithm.
 * Reference: code-templates/M_eq_log_M.c (4 square roots, Pade, rescale). */
void QLA_F3_M_veq_log_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_M_veq_log_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_F3_M_xeq_log_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_F3_M_xeq_log_pM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_F3_xM_eq_log_M(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_F3_M_eq_log_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_F3_xM_eq_log_xM(QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* F3: eigenvalues of a Hermitian matrix (nc = 3: closed form, template order, imaginary parts 0; QLA_FN nc = 4..8: the template's QR iteration).
 * Reference: code-templates/V_eq_eigenvalsH_M.c:103-161 (nc = 3), :180-484 (QR). */
void QLA_F3_V_veq_eigenvalsH_M(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_V_veq_eigenvalsH_pM(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_F3_V_xeq_eigenvalsH_M(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_F3_V_xeq_eigenvalsH_pM(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_F3_xV_eq_eigenvalsH_M(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_F3_V_eq_eigenvalsH_xM(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_F3_xV_eq_eigenvalsH_xM(QLA_F3_ColorVector *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_F3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);

/* D3: matrix exponential.
 * Reference: code-templates/M_eq_exp_M.c (scaled Pade(3,3), squaring). */
void QLA_D3_M_veq_exp_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_M_veq_exp_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_D3_M_xeq_exp_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_D3_M_xeq_exp_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_D3_xM_eq_exp_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_M_eq_exp_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_xM_eq_exp_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* D3: matrix square root.
 * Reference: code-templates/M_eq_sqrt_M.c (Newton iteration). */
void QLA_D3_M_veq_sqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_veq_sqrt_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_M_xeq_sqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_M_xeq_sqrt_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xM_eq_sqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_eq_sqrt_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xM_eq_sqrt_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);

/* D3: inverse square root.
 * Reference: code-templates/M_eq_invsqrt_M.c (Newton iteration). */
void QLA_D3_M_veq_invsqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_D3_M_veq_invsqrt_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int n);
void QLA_D3_M_xeq_invsqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *index,
                            int n);
void QLA_D3_M_xeq_invsqrt_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int *index,
                             int n);
void QLA_D3_xM_eq_invsqrt_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_D3_M_eq_invsqrt_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);
void QLA_D3_xM_eq_invsqrt_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             int n);

/* D3: determinant.
 * Reference: code-templates/C_eq_det_M.c (closed form for nc = 3, pivoted LU otherwise). */
void QLA_D3_C_veq_det_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_C_veq_det_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_D3_C_xeq_det_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_D3_C_xeq_det_pM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_D3_xC_eq_det_M(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_C_eq_det_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_xC_eq_det_xM(QLA_D_Complex *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* D3: matrix inverse.
 * Reference: code-templates/M_eq_inverse_M.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_D3_M_veq_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_D3_M_veq_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int n);
void QLA_D3_M_xeq_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *index,
                            int n);
void QLA_D3_M_xeq_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                             int *index,
                             int n);
void QLA_D3_xM_eq_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int n);
void QLA_D3_M_eq_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);
void QLA_D3_xM_eq_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                             int *indexr,
                             QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                             int *indexa,
                             int n);

/* D3: a^-1 b for a colour vector b.
 * Reference: code-templates/V_eq_M_inverse_V.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_D3_V_veq_M_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_V_veq_M_inverse_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_V_veq_pM_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_D3_V_veq_pM_inverse_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_D3_V_xeq_M_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_V_xeq_M_inverse_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_V_xeq_pM_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_D3_V_xeq_pM_inverse_pV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_D3_ColorVector *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_D3_V_eq_M_inverse_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_V_eq_xM_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_V_eq_xM_inverse_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_xV_eq_M_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_xV_eq_M_inverse_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_xV_eq_xM_inverse_V(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_D3_xV_eq_xM_inverse_xV(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                QLA_D3_ColorVector *QLA_MI355X_RESTRICT b,
                                int *indexb,
                                int n);

/* D3: a^-1 b for a colour matrix b.
 * Reference: code-templates/M_eq_M_inverse_M.c (Gauss-Jordan). */
void QLA_D3_M_veq_M_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_veq_M_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int n);
void QLA_D3_M_veq_pM_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_D3_M_veq_pM_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int n);
void QLA_D3_M_xeq_M_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *index,
                              int n);
void QLA_D3_M_xeq_M_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                               int *index,
                               int n);
void QLA_D3_M_xeq_pM_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *index,
                               int n);
void QLA_D3_M_xeq_pM_inverse_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *b,
                                int *index,
                                int n);
void QLA_D3_M_eq_M_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int *indexb,
                              int n);
void QLA_D3_M_eq_xM_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_M_eq_xM_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_xM_eq_M_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                              int n);
void QLA_D3_xM_eq_M_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int *indexb,
                               int n);
void QLA_D3_xM_eq_xM_inverse_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                               int n);
void QLA_D3_xM_eq_xM_inverse_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT b,
                                int *indexb,
                                int n);

/* D3: exponential of an anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_D3_M_veq_expA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_veq_expA_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int n);
void QLA_D3_M_xeq_expA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *index,
                         int n);
void QLA_D3_M_xeq_expA_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                          int *index,
                          int n);
void QLA_D3_xM_eq_expA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int n);
void QLA_D3_M_eq_expA_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);
void QLA_D3_xM_eq_expA_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);

/* D3: exponential of a traceless anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expTA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_D3_M_veq_expTA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_veq_expTA_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int n);
void QLA_D3_M_xeq_expTA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *index,
                          int n);
void QLA_D3_M_xeq_expTA_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                           int *index,
                           int n);
void QLA_D3_xM_eq_expTA_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          int *indexr,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int n);
void QLA_D3_M_eq_expTA_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                          QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                          int *indexa,
                          int n);
void QLA_D3_xM_eq_expTA_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);

/* D3: square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_sqrtPH_M.c (eigenvalue closed form for nc = 3, the sqrt branch otherwise). */
void QLA_D3_M_veq_sqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_D3_M_veq_sqrtPH_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int n);
void QLA_D3_M_xeq_sqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *index,
                           int n);
void QLA_D3_M_xeq_sqrtPH_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                            int *index,
                            int n);
void QLA_D3_xM_eq_sqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           int *indexr,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int n);
void QLA_D3_M_eq_sqrtPH_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                           QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                           int *indexa,
                           int n);
void QLA_D3_xM_eq_sqrtPH_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                            int *indexr,
                            QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                            int *indexa,
                            int n);

/* D3: inverse square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_invsqrtPH_M.c (closed form for nc = 3, the invsqrt branch otherwise). */
void QLA_D3_M_veq_invsqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_M_veq_invsqrtPH_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int n);
void QLA_D3_M_xeq_invsqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *index,
                              int n);
void QLA_D3_M_xeq_invsqrtPH_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                               int *index,
                               int n);
void QLA_D3_xM_eq_invsqrtPH_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              int *indexr,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int n);
void QLA_D3_M_eq_invsqrtPH_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                              QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                              int *indexa,
                              int n);
void QLA_D3_xM_eq_invsqrtPH_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);

/* D3: matrix logarithm.
 * Reference: code-templates/M_eq_log_M.c (4 square roots, Pade, rescale). */
void QLA_D3_M_veq_log_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_M_veq_log_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int n);
void QLA_D3_M_xeq_log_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *index,
                        int n);
void QLA_D3_M_xeq_log_pM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                         int *index,
                         int n);
void QLA_D3_xM_eq_log_M(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int n);
void QLA_D3_M_eq_log_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                        QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                        int *indexa,
                        int n);
void QLA_D3_xM_eq_log_xM(QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                         int *indexa,
                         int n);

/* D3: eigenvalues of a Hermitian matrix (nc = 3: closed form, template order, imaginary parts 0; QLA_FN nc = 4..8: the template's QR iteration).
 * Reference: code-templates/V_eq_eigenvalsH_M.c:103-161 (nc = 3), :180-484 (QR). */
void QLA_D3_V_veq_eigenvalsH_M(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_V_veq_eigenvalsH_pM(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int n);
void QLA_D3_V_xeq_eigenvalsH_M(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *index,
                               int n);
void QLA_D3_V_xeq_eigenvalsH_pM(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT *a,
                                int *index,
                                int n);
void QLA_D3_xV_eq_eigenvalsH_M(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               int *indexr,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int n);
void QLA_D3_V_eq_eigenvalsH_xM(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                               QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                               int *indexa,
                               int n);
void QLA_D3_xV_eq_eigenvalsH_xM(QLA_D3_ColorVector *QLA_MI355X_RESTRICT r,
                                int *indexr,
                                QLA_D3_ColorMatrix *QLA_MI355X_RESTRICT a,
                                int *indexa,
                                int n);

/* FN: matrix exponential.
 * Reference: code-templates/M_eq_exp_M.c (scaled Pade(3,3), squaring). */
void QLA_FN_M_veq_exp_M(int nc,
                        void *r,
                        void *a,
                        int n);
void QLA_FN_M_veq_exp_pM(int nc,
                         void *r,
                         void *a,
                         int n);
void QLA_FN_M_xeq_exp_M(int nc,
                        void *r,
                        void *a,
                        int *index,
                        int n);
void QLA_FN_M_xeq_exp_pM(int nc,
                         void *r,
                         void *a,
                         int *index,
                         int n);
void QLA_FN_xM_eq_exp_M(int nc,
                        void *r,
                        int *indexr,
                        void *a,
                        int n);
void QLA_FN_M_eq_exp_xM(int nc,
                        void *r,
                        void *a,
                        int *indexa,
                        int n);
void QLA_FN_xM_eq_exp_xM(int nc,
                         void *r,
                         int *indexr,
                         void *a,
                         int *indexa,
                         int n);

/* FN: matrix square root.
 * Reference: code-templates/M_eq_sqrt_M.c (Newton iteration). */
void QLA_FN_M_veq_sqrt_M(int nc,
                         void *r,
                         void *a,
                         int n);
void QLA_FN_M_veq_sqrt_pM(int nc,
                          void *r,
                          void *a,
                          int n);
void QLA_FN_M_xeq_sqrt_M(int nc,
                         void *r,
                         void *a,
                         int *index,
                         int n);
void QLA_FN_M_xeq_sqrt_pM(int nc,
                          void *r,
                          void *a,
                          int *index,
                          int n);
void QLA_FN_xM_eq_sqrt_M(int nc,
                         void *r,
                         int *indexr,
                         void *a,
                         int n);
void QLA_FN_M_eq_sqrt_xM(int nc,
                         void *r,
                         void *a,
                         int *indexa,
                         int n);
void QLA_FN_xM_eq_sqrt_xM(int nc,
                          void *r,
                          int *indexr,
                          void *a,
                          int *indexa,
                          int n);

/* FN: inverse square root.
 * Reference: code-templates/M_eq_invsqrt_M.c (Newton iteration). */
void QLA_FN_M_veq_invsqrt_M(int nc,
                            void *r,
                            void *a,
                            int n);
void QLA_FN_M_veq_invsqrt_pM(int nc,
                             void *r,
                             void *a,
                             int n);
void QLA_FN_M_xeq_invsqrt_M(int nc,
                            void *r,
                            void *a,
                            int *index,
                            int n);
void QLA_FN_M_xeq_invsqrt_pM(int nc,
                             void *r,
                             void *a,
                             int *index,
                             int n);
void QLA_FN_xM_eq_invsqrt_M(int nc,
                            void *r,
                            int *indexr,
                            void *a,
                            int n);
void QLA_FN_M_eq_invsqrt_xM(int nc,
                            void *r,
                            void *a,
                            int *indexa,
                            int n);
void QLA_FN_xM_eq_invsqrt_xM(int nc,
                             void *r,
                             int *indexr,
                             void *a,
                             int *indexa,
                             int n);

/* FN: determinant.
 * Reference: code-templates/C_eq_det_M.c (closed form for nc = 3, pivoted LU otherwise). */
void QLA_FN_C_veq_det_M(int nc,
                        QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        void *a,
                        int n);
void QLA_FN_C_veq_det_pM(int nc,
                         QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         void *a,
                         int n);
void QLA_FN_C_xeq_det_M(int nc,
                        QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        void *a,
                        int *index,
                        int n);
void QLA_FN_C_xeq_det_pM(int nc,
                         QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         void *a,
                         int *index,
                         int n);
void QLA_FN_xC_eq_det_M(int nc,
                        QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        int *indexr,
                        void *a,
                        int n);
void QLA_FN_C_eq_det_xM(int nc,
                        QLA_F_Complex *QLA_MI355X_RESTRICT r,
                        void *a,
                        int *indexa,
                        int n);
void QLA_FN_xC_eq_det_xM(int nc,
                         QLA_F_Complex *QLA_MI355X_RESTRICT r,
                         int *indexr,
                         void *a,
                         int *indexa,
                         int n);

/* FN: matrix inverse.
 * Reference: code-templates/M_eq_inverse_M.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_FN_M_veq_inverse_M(int nc,
                            void *r,
                            void *a,
                            int n);
void QLA_FN_M_veq_inverse_pM(int nc,
                             void *r,
                             void *a,
                             int n);
void QLA_FN_M_xeq_inverse_M(int nc,
                            void *r,
                            void *a,
                            int *index,
                            int n);
void QLA_FN_M_xeq_inverse_pM(int nc,
                             void *r,
                             void *a,
                             int *index,
                             int n);
void QLA_FN_xM_eq_inverse_M(int nc,
                            void *r,
                            int *indexr,
                            void *a,
                            int n);
void QLA_FN_M_eq_inverse_xM(int nc,
                            void *r,
                            void *a,
                            int *indexa,
                            int n);
void QLA_FN_xM_eq_inverse_xM(int nc,
                             void *r,
                             int *indexr,
                             void *a,
                             int *indexa,
                             int n);

/* FN: a^-1 b for a colour vector b.
 * Reference: code-templates/V_eq_M_inverse_V.c (adjugate for nc = 3, Gauss-Jordan otherwise). */
void QLA_FN_V_veq_M_inverse_V(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int n);
void QLA_FN_V_veq_M_inverse_pV(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int n);
void QLA_FN_V_veq_pM_inverse_V(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int n);
void QLA_FN_V_veq_pM_inverse_pV(int nc,
                                void *r,
                                void *a,
                                void *b,
                                int n);
void QLA_FN_V_xeq_M_inverse_V(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int *index,
                              int n);
void QLA_FN_V_xeq_M_inverse_pV(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int *index,
                               int n);
void QLA_FN_V_xeq_pM_inverse_V(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int *index,
                               int n);
void QLA_FN_V_xeq_pM_inverse_pV(int nc,
                                void *r,
                                void *a,
                                void *b,
                                int *index,
                                int n);
void QLA_FN_V_eq_M_inverse_xV(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int *indexb,
                              int n);
void QLA_FN_V_eq_xM_inverse_V(int nc,
                              void *r,
                              void *a,
                              int *indexa,
                              void *b,
                              int n);
void QLA_FN_V_eq_xM_inverse_xV(int nc,
                               void *r,
                               void *a,
                               int *indexa,
                               void *b,
                               int *indexb,
                               int n);
void QLA_FN_xV_eq_M_inverse_V(int nc,
                              void *r,
                              int *indexr,
                              void *a,
                              void *b,
                              int n);
void QLA_FN_xV_eq_M_inverse_xV(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               void *b,
                               int *indexb,
                               int n);
void QLA_FN_xV_eq_xM_inverse_V(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               int *indexa,
                               void *b,
                               int n);
void QLA_FN_xV_eq_xM_inverse_xV(int nc,
                                void *r,
                                int *indexr,
                                void *a,
                                int *indexa,
                                void *b,
                                int *indexb,
                                int n);

/* FN: a^-1 b for a colour matrix b.
 * Reference: code-templates/M_eq_M_inverse_M.c (Gauss-Jordan). */
void QLA_FN_M_veq_M_inverse_M(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int n);
void QLA_FN_M_veq_M_inverse_pM(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int n);
void QLA_FN_M_veq_pM_inverse_M(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int n);
void QLA_FN_M_veq_pM_inverse_pM(int nc,
                                void *r,
                                void *a,
                                void *b,
                                int n);
void QLA_FN_M_xeq_M_inverse_M(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int *index,
                              int n);
void QLA_FN_M_xeq_M_inverse_pM(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int *index,
                               int n);
void QLA_FN_M_xeq_pM_inverse_M(int nc,
                               void *r,
                               void *a,
                               void *b,
                               int *index,
                               int n);
void QLA_FN_M_xeq_pM_inverse_pM(int nc,
                                void *r,
                                void *a,
                                void *b,
                                int *index,
                                int n);
void QLA_FN_M_eq_M_inverse_xM(int nc,
                              void *r,
                              void *a,
                              void *b,
                              int *indexb,
                              int n);
void QLA_FN_M_eq_xM_inverse_M(int nc,
                              void *r,
                              void *a,
                              int *indexa,
                              void *b,
                              int n);
void QLA_FN_M_eq_xM_inverse_xM(int nc,
                               void *r,
                               void *a,
                               int *indexa,
                               void *b,
                               int *indexb,
                               int n);
void QLA_FN_xM_eq_M_inverse_M(int nc,
                              void *r,
                              int *indexr,
                              void *a,
                              void *b,
                              int n);
void QLA_FN_xM_eq_M_inverse_xM(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               void *b,
                               int *indexb,
                               int n);
void QLA_FN_xM_eq_xM_inverse_M(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               int *indexa,
                               void *b,
                               int n);
void QLA_FN_xM_eq_xM_inverse_xM(int nc,
                                void *r,
                                int *indexr,
                                void *a,
                                int *indexa,
                                void *b,
                                int *indexb,
                                int n);

/* FN: exponential of an anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_FN_M_veq_expA_M(int nc,
                         void *r,
                         void *a,
                         int n);
void QLA_FN_M_veq_expA_pM(int nc,
                          void *r,
                          void *a,
                          int n);
void QLA_FN_M_xeq_expA_M(int nc,
                         void *r,
                         void *a,
                         int *index,
                         int n);
void QLA_FN_M_xeq_expA_pM(int nc,
                          void *r,
                          void *a,
                          int *index,
                          int n);
void QLA_FN_xM_eq_expA_M(int nc,
                         void *r,
                         int *indexr,
                         void *a,
                         int n);
void QLA_FN_M_eq_expA_xM(int nc,
                         void *r,
                         void *a,
                         int *indexa,
                         int n);
void QLA_FN_xM_eq_expA_xM(int nc,
                          void *r,
                          int *indexr,
                          void *a,
                          int *indexa,
                          int n);

/* FN: exponential of a traceless anti-Hermitian matrix.
 * Reference: code-templates/M_eq_expTA_M.c (Cayley-Hamilton for nc = 3, the exp branch otherwise). */
void QLA_FN_M_veq_expTA_M(int nc,
                          void *r,
                          void *a,
                          int n);
void QLA_FN_M_veq_expTA_pM(int nc,
                           void *r,
                           void *a,
                           int n);
void QLA_FN_M_xeq_expTA_M(int nc,
                          void *r,
                          void *a,
                          int *index,
                          int n);
void QLA_FN_M_xeq_expTA_pM(int nc,
                           void *r,
                           void *a,
                           int *index,
                           int n);
void QLA_FN_xM_eq_expTA_M(int nc,
                          void *r,
                          int *indexr,
                          void *a,
                          int n);
void QLA_FN_M_eq_expTA_xM(int nc,
                          void *r,
                          void *a,
                          int *indexa,
                          int n);
void QLA_FN_xM_eq_expTA_xM(int nc,
                           void *r,
                           int *indexr,
                           void *a,
                           int *indexa,
                           int n);

/* FN: square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_sqrtPH_M.c (eigenvalue closed form for nc = 3, the sqrt branch otherwise). */
void QLA_FN_M_veq_sqrtPH_M(int nc,
                           void *r,
                           void *a,
                           int n);
void QLA_FN_M_veq_sqrtPH_pM(int nc,
                            void *r,
                            void *a,
                            int n);
void QLA_FN_M_xeq_sqrtPH_M(int nc,
                           void *r,
                           void *a,
                           int *index,
                           int n);
void QLA_FN_M_xeq_sqrtPH_pM(int nc,
                            void *r,
                            void *a,
                            int *index,
                            int n);
void QLA_FN_xM_eq_sqrtPH_M(int nc,
                           void *r,
                           int *indexr,
                           void *a,
                           int n);
void QLA_FN_M_eq_sqrtPH_xM(int nc,
                           void *r,
                           void *a,
                           int *indexa,
                           int n);
void QLA_FN_xM_eq_sqrtPH_xM(int nc,
                            void *r,
                            int *indexr,
                            void *a,
                            int *indexa,
                            int n);

/* FN: inverse square root of a positive Hermitian matrix.
 * Reference: code-templates/M_eq_invsqrtPH_M.c (closed form for nc = 3, the invsqrt branch otherwise). */
void QLA_FN_M_veq_invsqrtPH_M(int nc,
                              void *r,
                              void *a,
                              int n);
void QLA_FN_M_veq_invsqrtPH_pM(int nc,
                               void *r,
                               void *a,
                               int n);
void QLA_FN_M_xeq_invsqrtPH_M(int nc,
                              void *r,
                              void *a,
                              int *index,
                              int n);
void QLA_FN_M_xeq_invsqrtPH_pM(int nc,
                               void *r,
                               void *a,
                               int *index,
                               int n);
void QLA_FN_xM_eq_invsqrtPH_M(int nc,
                              void *r,
                              int *indexr,
                              void *a,
                              int n);
void QLA_FN_M_eq_invsqrtPH_xM(int nc,
                              void *r,
                              void *a,
                              int *indexa,
                              int n);
void QLA_FN_xM_eq_invsqrtPH_xM(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               int *indexa,
                               int n);

/* FN: matrix logarithm.
 * Reference: code-templates/M_eq_log_M.c (4 square roots, Pade, rescale). */
void QLA_FN_M_veq_log_M(int nc,
                        void *r,
                        void *a,
                        int n);
void QLA_FN_M_veq_log_pM(int nc,
                         void *r,
                         void *a,
                         int n);
void QLA_FN_M_xeq_log_M(int nc,
                        void *r,
                        void *a,
                        int *index,
                        int n);
void QLA_FN_M_xeq_log_pM(int nc,
                         void *r,
                         void *a,
                         int *index,
                         int n);
void QLA_FN_xM_eq_log_M(int nc,
                        void *r,
                        int *indexr,
                        void *a,
                        int n);
void QLA_FN_M_eq_log_xM(int nc,
                        void *r,
                        void *a,
                        int *indexa,
                        int n);
void QLA_FN_xM_eq_log_xM(int nc,
                         void *r,
                         int *indexr,
                         void *a,
                         int *indexa,
                         int n);

/* FN: eigenvalues of a Hermitian matrix (nc = 3: closed form, template order, imaginary parts 0; QLA_FN nc = 4..8: the template's QR iteration).
 * Reference: code-templates/V_eq_eigenvalsH_M.c:103-161 (nc = 3), :180-484 (QR). */
void QLA_FN_V_veq_eigenvalsH_M(int nc,
                               void *r,
                               void *a,
                               int n);
void QLA_FN_V_veq_eigenvalsH_pM(int nc,
                                void *r,
                                void *a,
                                int n);
void QLA_FN_V_xeq_eigenvalsH_M(int nc,
                               void *r,
                               void *a,
                               int *index,
                               int n);
void QLA_FN_V_xeq_eigenvalsH_pM(int nc,
                                void *r,
                                void *a,
                                int *index,
                                int n);
void QLA_FN_xV_eq_eigenvalsH_M(int nc,
                               void *r,
                               int *indexr,
                               void *a,
                               int n);
void QLA_FN_V_eq_eigenvalsH_xM(int nc,
                               void *r,
                               void *a,
                               int *indexa,
                               int n);
void QLA_FN_xV_eq_eigenvalsH_xM(int nc,
                                void *r,
                                int *indexr,
                                void *a,
                                int *indexa,
                                int n);

#endif
