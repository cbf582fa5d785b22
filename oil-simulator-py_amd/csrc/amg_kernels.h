// Launchers of the Ruge-Stueben setup kernels (amg_setup.hip) and of the dense
// coarsest-level factor/solve (coarse.hip).
#pragma once

#include "common.h"

namespace impes {

constexpr signed char kC = 0;  // coarse point marker (reference: cf_marker 0 = C)
constexpr signed char kF = 1;  // fine point marker

// diag[i] = a_ii (0 if absent); inv_relax = min(1/max(|a_ii|, 0.3*sum_j|a_ij|, 1e-30), 1e2)
void launch_diag_relax(int n, const int* rp, const int* ci, const double* v, double* diag, double* inv_relax64,
                       float* inv_relax32, hipStream_t s);
// launch_diag_relax (diag not written) and the padded FP32 values v4 of the same
// rows (as launch_csr4_values) in one pass
void launch_diag_relax_pad4(int n, const int* rp, const int* ci, const double* v, const int* rp4, float* v4,
                            double* ir64, float* ir32, hipStream_t s);
// number of rows with |a_ii| < thr (lower-median test of the equilibration switch)
void launch_count_small_diag(int n, const double* diag, double thr, int* flags, hipStream_t s);
// dh[i] = 1/sqrt(max(|diag_i|, 1e-30))
void launch_dhalf_inv(int n, const double* diag, double* dh, hipStream_t s);
// vo[e] = v[e]*dh[row]*dh[col]
void launch_sym_scale(int n, const int* rp, const int* ci, const double* v, const double* dh, double* vo,
                      hipStream_t s);

void launch_strength(int n, long long nnz, const int* rp, const int* ci, const double* v, double theta,
                     unsigned char* strong,
                     hipStream_t s);
// tpos[e] = position of (col,row) in row col, -1 if absent
void launch_transpose_pos(int n, long long nnz, const int* rp, const int* ci, int* tpos, hipStream_t s);
// C/F split: symmetrised strong graph, strict local maxima of (deg, hash, index), lift of F points with no C neighbour
void launch_cf_split(int n, long long nnz, const int* rp, const int* ci, const unsigned char* strong,
                     const int* tpos, int* deg, signed char* is_max, signed char* cf, hipStream_t s);
// iterated independent-set split on the symmetrised strong graph with the same
// (deg, hash, index) order: rounds of local-max selection among undecided points,
// whose undecided neighbours become F, until every point is decided (no lift needed)
void launch_cf_split_iterated(int n, long long nnz, const int* rp, const int* ci, const unsigned char* strong,
                              const int* tpos, int* deg, signed char* newc, int* undecided, signed char* cf,
                              hipStream_t s);
// cf[i] = C where |a_ii| > ratio * sum_j!=i |a_ij| (Dirichlet-like rows)
void launch_force_dominant(int n, const int* rp, const int* ci, const double* v, double ratio, signed char* cf,
                           hipStream_t s);
// promote F points with no strong C neighbour
void launch_orphans(int n, const int* rp, const int* ci, const unsigned char* strong, const signed char* cf_in,
                    signed char* cf_out, int* is_c, hipStream_t s);

// interpolation
void launch_interp_pass_a(int n, const int* rp, const int* ci, const double* v, const unsigned char* strong,
                          const signed char* cf, double* sum_strong, int* cnt_c, hipStream_t s);
void launch_interp_count(int n, const int* rp, const int* ci, const unsigned char* strong, const signed char* cf,
                         const double* sum_strong, const int* cnt_c, int* cand, hipStream_t s);
void launch_interp_expand(int n, long long nnz, const int* rp, const int* ci, const double* v,
                          const unsigned char* strong, const signed char* cf, const double* sum_strong,
                          const int* cnt_c, const int* f2c, const long long* cand_ptr, unsigned long long* keys,
                          double* vals, hipStream_t s);
// final F-row weights; trunc > 0 / pmax > 0 truncate each F row afterwards: weights below
// trunc * row max are dropped, then all but the pmax largest (ties: lower coarse column
// first), and the kept weights are rescaled to the row's previous sum
void launch_interp_reduce(int n, const int* rp, const int* ci, const signed char* cf, const int* f2c,
                          const long long* cand_ptr, const unsigned long long* keys, const double* vals,
                          double trunc, int pmax, int* out_cols, double* out_vals, int* pcnt, hipStream_t s);
void launch_interp_gather(int n, const long long* cand_ptr, const int* cols_tmp, const double* vals_tmp,
                          const int* prp, int* pci, double* pv, hipStream_t s);

// transpose (R = P^T): counts per column, fill, key build, gather after sort
void launch_col_count(int nnz, const int* ci, int* cnt, hipStream_t s);
void launch_transpose_fill(int n_rows, const int* rp, const int* ci, const int* rrp, int* cursor,
                           unsigned long long* keys, hipStream_t s);
// after the per-column sort: rci[p] = source row, rv[p] = source value
void launch_transpose_finish(int nnz, const unsigned long long* keys, const double* src_v, int* rci, double* rv,
                             hipStream_t s);

// SpGEMM C = L*M, symbolic through expanded products sorted per row
void launch_spgemm_count(int n, const int* lrp, const int* lci, const int* mrp, int* cnt, hipStream_t s);
// keys[t] = (column, local index); ex_e/ex_f[t] = (left entry, right entry) in expansion order
void launch_spgemm_expand(int n, long long left_nnz, long long products, const int* lrp, const int* lci, const int* mrp, const int* mci,
                          const long long* prod_ptr, unsigned long long* keys, int* ex_e, int* ex_f, hipStream_t s);
void launch_spgemm_unique_count(int n, const long long* prod_ptr, const unsigned long long* keys, int* cnt,
                                hipStream_t s);
// writes C column indices, product segment pointers (int64, nnzC+1) and the sorted
// operand index pairs (left entry, right entry) of every product
void launch_spgemm_finalize(int n, const long long* prod_ptr, const unsigned long long* keys, const int* crp,
                            const int* ex_e, const int* ex_f, int* cci, long long* seg, int* pe, int* pf,
                            long long total_products, int nnz_c, hipStream_t s);
// C.v[o] = sum_{t in seg(o)} Lv[pe[t]] * Mv[pf[t]]
// lanes per output entry of the numeric product from its number of terms
int spgemm_group_for(long long terms, int nnz_c);
void launch_spgemm_numeric(int nnz_c, const long long* seg, const int* pe, const int* pf, const double* lv,
                           const double* mv, double* cv, int group, hipStream_t s);

// coarsest level (coarse.hip). pin is a device flag written by launch_pin_decide and
// read by the kernels below, so the decision never leaves the device.
// pin = (mode == reference) || (mode == auto && every row has |row sum| <= 1e-8 |a_ii|)
void launch_pin_decide(int n, const int* rp, const int* ci, const double* v, int mode, int* pin, hipStream_t s);
// dense inverse of the (optionally DOF-0-pinned) matrix
void launch_csr_to_dense(int n, const int* rp, const int* ci, const double* v, const int* pin, double* dense,
                         hipStream_t s);
// aug is n x 2n row-major [A | I], inverse into the right half: the LDS sweep for symmetric
// positive definite A (n <= 199), else pivoted Gauss-Jordan; done: device int scratch
void launch_gauss_jordan_inverse(int n, double* aug, int* done, hipStream_t s);
// x = Ainv * b (with pin: b[0] treated as 0); FP64 and FP32 vector variants
// Coarsest levels of more than kDenseSweepMax rows (up to kDenseMfmaMax): the inverse by
// blocked Gauss-Jordan on the matrix cores (dense_mfma.hip) into M (dense_pad(n)^2
// doubles), with W (dense_pad(n)^2) and P (2 x 64 x 64) as work; *bad is set when a
// pivot is not positive (the operator is not SPD). Applied by launch_dense_mv_*.
constexpr int kDenseMfmaMax = 4096;
constexpr int kDenseSweepMax = 160;  // coarsest levels up to this size: the register sweep
int dense_pad(int n);
void launch_dense_inverse_mfma(int n, const int* rp, const int* ci, const double* v, const int* pin, double* M,
                               double* W, double* P, int* bad, hipStream_t s);
void launch_dense_mv_f64(int n, const int* pin, const double* M, const double* b, double* x, hipStream_t s);
void launch_dense_mv_f32(int n, const int* pin, const double* M, const float* b, float* x, hipStream_t s);
void launch_dense_solve_f64(int n, const int* pin, const double* aug, const double* b, double* x, hipStream_t s);
void launch_dense_solve_f32(int n, const int* pin, const double* aug, const float* b, float* x, hipStream_t s);
// b[0] = 0 when *pin (Jacobi coarsest path)
void launch_pin_rhs_f64(const int* pin, double* b, hipStream_t s);
void launch_pin_rhs_f32(const int* pin, float* b, hipStream_t s);

}  // namespace impes
