// Host-side launchers of the HIP kernels (one translation unit per kernel family).
// All launchers enqueue on the given stream and never synchronise.
#pragma once

#include <vector>

#include "common.h"

namespace impes {

// ------------------------------------------------------------------ blas.hip
constexpr int kRedBlocks = 512;  // fixed partial count of two-stage reductions
constexpr int kMaxMultiDot = 128;
constexpr int kMultiDotBlocks = 2048;  // partials per vector of launch_multidot_f64 (work: k * 2048 doubles)

void launch_fill_f64(double* x, double v, long long n, hipStream_t s);
void launch_fill_f32(float* x, float v, long long n, hipStream_t s);
void launch_fill_i32(int* x, int v, long long n, hipStream_t s);
void launch_copy_f64(double* y, const double* x, long long n, hipStream_t s);
void launch_copy_f32(float* y, const float* x, long long n, hipStream_t s);
void launch_f64_to_f32(float* y, const double* x, long long n, hipStream_t s);
void launch_f32_to_f64(double* y, const float* x, long long n, hipStream_t s);
// y = a .* x (elementwise), converting precision as needed
void launch_scale_f64_to_f32(float* y, const double* a, const double* x, long long n, hipStream_t s);
void launch_scale_f32_to_f64(double* y, const double* a, const float* x, long long n, hipStream_t s);
void launch_scale_f64(double* y, const double* a, const double* x, long long n, hipStream_t s);
// y += alpha * x  (alpha host scalar)
void launch_axpy_f64(double* y, double alpha, const double* x, long long n, hipStream_t s);
// y = x ./ a
void launch_div_f64(double* y, const double* x, const double* a, long long n, hipStream_t s);
// y = x + beta * y
void launch_xpby_f64(double* y, const double* x, double beta, long long n, hipStream_t s);
// part[b] (b < nb) = block partial sums of x . y (nb blocks of a grid-stride loop)
void launch_dot_partial_f64(const double* x, const double* y, long long n, double* part, int nb, hipStream_t s);
// out[0] = sum(x .* y); work >= kRedBlocks doubles
void launch_dot_f64(const double* x, const double* y, long long n, double* out, double* work, hipStream_t s);
// out[0] = max |x|
void launch_amax_f64(const double* x, long long n, double* out, double* work, hipStream_t s);
// h[i] = V_i . w  for i < k   (V_i = V + i*ldv); work >= kMultiDotBlocks*k doubles
void launch_multidot_f64(const double* V, long long ldv, int k, const double* w, long long n,
                         double* h, double* work, hipStream_t s);
// w -= sum_i V_i h[i]
void launch_gs_update_f64(double* w, const double* V, long long ldv, int k, const double* h,
                          long long n, hipStream_t s);
// v_out = w / nrm[0]   (nrm device scalar)
void launch_scale_by_inv_f64(double* v_out, const double* w, const double* nrm, long long n, hipStream_t s);
// x += sum_i Z_i y[i]
void launch_lincomb_add_f64(double* x, const double* Z, long long ldz, int k, const double* y,
                            long long n, hipStream_t s);

// ------------------------------------------------------------------ spmv.hip
struct CsrF64 { int n_rows = 0, n_cols = 0, nnz = 0; const int* rp = nullptr; const int* ci = nullptr; const double* v = nullptr; };
struct CsrF32 { int n_rows = 0, n_cols = 0, nnz = 0; const int* rp = nullptr; const int* ci = nullptr; const float* v = nullptr; };

// Structured 7-point symmetric operator on an (nx,ny,nz) grid, z fastest:
// A[c,c] = d[c]; A[c,c+sx] = A[c+sx,c] = ox[c]; A[c,c+nz] = oy[c]; A[c,c+1] = oz[c].
// ox/oy/oz hold 0 on the faces that leave the grid.
template <class T>
struct Stencil7 { int nx = 0, ny = 0, nz = 0; const T* d = nullptr; const T* ox = nullptr; const T* oy = nullptr; const T* oz = nullptr; };

// FP32 CSR with every row padded to a multiple of 4 entries (value 0, column = the
// row itself), 16-byte aligned: a lane reads 4 consecutive entries of its row with
// one vector load, and the lanes of a wavefront read whole cache lines whatever
// the row lengths. rp4 counts entries (multiples of 4); nnz is the unpadded count
// (algorithmic bytes).
struct Csr4F32 {
  int n = 0, nnz = 0;
  const int* rp4 = nullptr;
  const int4* ci4 = nullptr;
  const float4* v4 = nullptr;
  // when set (every |column - row| < 2^15): 16-bit column offsets from the row,
  // read instead of ci4 (8 bytes per 4 entries instead of 16)
  const short4* co4 = nullptr;
  // when set: the rows cut into nw wave tiles of whole rows and at most 64 chunks
  // (tile t = rows wrow[t] .. wrow[t+1]-1, chunks wq[t] .. wq[t+1]-1)
  int nw = 0;
  const int* wrow = nullptr;
  const int* wq = nullptr;
};
// Wave tiles of a padded operator (see Csr4F32): rows packed greedily, in order,
// into tiles of at most 64 four-entry chunks. Fills wrow and wq (nw + 1 entries
// each) from the host copy of rp4 (n + 1 entry offsets); returns nw, or 0 when a
// row has more than 64 chunks.
int csr4_wave_tiles(int n, const int* rp4_host, std::vector<int>& wrow, std::vector<int>& wq);
// lanes per row of the padded kernels from the padded entry count
int csr4_group_for(int n_rows, long long nnz4);
// padded row lengths: cnt[r] = rows' lengths rounded up to 4 (an empty row gets 4)
void launch_csr4_count(int n, const int* rp, int* cnt, hipStream_t s);
// padded pattern (ci4 sized rp4[n]) and FP32 values (rounded from FP64) of a CSR
// matrix; padding entries have value 0 and the column of the row itself (square)
// or of the row's first entry (rectangular; 0 for an empty row)
void launch_csr4_pattern(int n, const int* rp, const int* ci, const int* rp4, int* ci4, hipStream_t s,
                         bool square = true);
void launch_csr4_values(int n, const int* rp, const double* v, const int* rp4, float* v4, hipStream_t s);
// co4[e] = ci4[e] - row (16-bit) over nnz4 entries; *maxabs = the largest |ci4[e] - row|
// (a device int, zero on entry)
void launch_csr4_offsets(int n, const int* rp4, const int* ci4, short* co4, int* maxabs, hipStream_t s);
// r = b - A x and xo = xi + w inv_relax .* (b - A xi) on the padded form
void launch_residual4_f32(const Csr4F32& A, int group, const float* x, const float* b, float* r, hipStream_t s);
void launch_jacobi4_f32(const Csr4F32& A, int group, const float* xi, const float* b, const float* inv_relax, float w,
                        float* xo, hipStream_t s);
// restriction on a padded operator with wave tiles (A.wrow set; column indices in
// ci4): y = A x; and y = A x with y2 = w inv_relax2 .* y (fused with the coarse
// level's first sweep from zero)
void launch_spmv4_f32(const Csr4F32& A, const float* x, float* y, hipStream_t s);
void launch_restrict4_j0_f32(const Csr4F32& R, const float* r, const float* inv_relax_c, float w, float* bc, float* xc,
                             hipStream_t s);

// pick lanes-per-row from the mean row length
int csr_group_for(int n_rows, int nnz);

void launch_spmv_f64(const CsrF64& A, const double* x, double* y, int group, hipStream_t s);
void launch_spmv_f32(const CsrF32& A, const float* x, float* y, int group, hipStream_t s);
// y += A x
void launch_spmv_add_f64(const CsrF64& A, const double* x, double* y, int group, hipStream_t s);
void launch_spmv_add_f32(const CsrF32& A, const float* x, float* y, int group, hipStream_t s);
// y = y0 + A x  (y0 and y distinct)
void launch_spmv_add_to_f64(const CsrF64& A, const double* x, const double* y0, double* y, int group, hipStream_t s);
void launch_spmv_add_to_f32(const CsrF32& A, const float* x, const float* y0, float* y, int group, hipStream_t s);
// r = b - A x
void launch_residual_f64(const CsrF64& A, const double* x, const double* b, double* r, int group, hipStream_t s);
void launch_residual_f32(const CsrF32& A, const float* x, const float* b, float* r, int group, hipStream_t s);
// xo = xi + w * inv_relax .* (b - A xi)
void launch_jacobi_f64(const CsrF64& A, const double* xi, const double* b, const double* inv_relax,
                       double w, double* xo, int group, hipStream_t s);
void launch_jacobi_f32(const CsrF32& A, const float* xi, const float* b, const float* inv_relax,
                       float w, float* xo, int group, hipStream_t s);
// xo = w * inv_relax .* b  (first sweep from a zero guess)
// bc = R r and xc = w .* inv_relax_c .* bc (restriction fused with the coarse
// level's first L1-Jacobi sweep from a zero guess)
void launch_restrict_j0_f64(const CsrF64& R, const double* r, const double* inv_relax_c, double w, double* bc,
                            double* xc, int group, hipStream_t s);
void launch_restrict_j0_f32(const CsrF32& R, const float* r, const float* inv_relax_c, float w, float* bc, float* xc,
                            int group, hipStream_t s);
void launch_jacobi0_f64(const double* b, const double* inv_relax, double w, double* xo, int n, hipStream_t s);
void launch_jacobi0_f32(const float* b, const float* inv_relax, float w, float* xo, int n, hipStream_t s);

void launch_st7_spmv_f64(const Stencil7<double>& S, const double* x, double* y, hipStream_t s);
// r = b - A x
void launch_st7_residual_f64(const Stencil7<double>& S, const double* x, const double* b, double* r, hipStream_t s);
void launch_st7_residual_f32(const Stencil7<float>& S, const float* x, const float* b, float* r, hipStream_t s);
void launch_st7_jacobi_f64(const Stencil7<double>& S, const double* xi, const double* b, const double* inv_relax,
                           double w, double* xo, hipStream_t s);
void launch_st7_jacobi_f32(const Stencil7<float>& S, const float* xi, const float* b, const float* inv_relax,
                           float w, float* xo, hipStream_t s);
// two L1-Jacobi sweeps from a zero guess in one pass: with x1 = w * inv_relax .* b,
// xo = x1 + w * inv_relax .* (b - A x1); equal to launch_jacobi0 then launch_st7_jacobi
void launch_st7_jacobi_z0_f64(const Stencil7<double>& S, const double* b, const double* inv_relax, double w,
                              double* xo, hipStream_t s);
void launch_st7_jacobi_z0_f32(const Stencil7<float>& S, const float* b, const float* inv_relax, float w, float* xo,
                              hipStream_t s);
// y = dinv .* (A x)   (row-scaled operator used by the outer Krylov method)
void launch_st7_rowscaled_spmv_f64(const Stencil7<double>& S, const double* dinv, const double* x, double* y, hipStream_t s);
// z = sout .* x (x the FP32 or FP64 V-cycle output) and y = dinv .* (A z) in one pass
void launch_st7_rowscaled_spmv_z_f64(const Stencil7<double>& S, const double* dinv, const double* sout,
                                     const float* x, double* z, double* y, hipStream_t s);
void launch_st7_rowscaled_spmv_z_f64(const Stencil7<double>& S, const double* dinv, const double* sout,
                                     const double* x, double* z, double* y, hipStream_t s);
void launch_csr_rowscaled_spmv_f64(const CsrF64& A, const double* dinv, const double* x, double* y, int group, hipStream_t s);
// r = dinv .* (b - A x)
void launch_st7_rowscaled_residual_f64(const Stencil7<double>& S, const double* dinv, const double* x,
                                       const double* b, double* r, hipStream_t s);
void launch_csr_rowscaled_residual_f64(const CsrF64& A, const double* dinv, const double* x, const double* b,
                                       double* r, int group, hipStream_t s);

}  // namespace impes
