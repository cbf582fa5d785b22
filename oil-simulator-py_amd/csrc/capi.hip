// extern "C" entry points: argument checking, exception -> error code mapping.
#include <string>

#include "../../include/impes_hip.h"
#include "amg.h"
#include "amg_kernels.h"
#include "fgmres.h"

namespace impes {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace impes

using namespace impes;

struct impes_amg {
  Hierarchy h;
  explicit impes_amg(const AmgParams& p) : h(p) {}
};
struct impes_solver {
  Fgmres f;
};

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static void check_csr(int n_rows, int nnz, const int* rowptr, const int* col, const void* val) {
  IMPES_REQUIRE(n_rows >= 0 && nnz >= 0, "negative size");
  IMPES_REQUIRE(rowptr && (nnz == 0 || (col && val)), "null CSR pointer");
}

extern "C" {

const char* impes_last_error(void) { return g_last_error.c_str(); }
int impes_abi_version(void) { return IMPES_ABI_VERSION; }

int impes_spmv_csr_f64(int n_rows, int n_cols, int nnz, const int* rowptr, const int* col, const double* val,
                       const double* x, double* y, void* stream) {
  return guarded(__func__, [&] {
    check_csr(n_rows, nnz, rowptr, col, val);
    IMPES_REQUIRE(x && y, "null vector");
    CsrF64 A{n_rows, n_cols, nnz, rowptr, col, val};
    launch_spmv_f64(A, x, y, csr_group_for(n_rows, nnz), as_stream(stream));
  });
}

int impes_spmv_csr_f32(int n_rows, int n_cols, int nnz, const int* rowptr, const int* col, const float* val,
                       const float* x, float* y, void* stream) {
  return guarded(__func__, [&] {
    check_csr(n_rows, nnz, rowptr, col, val);
    IMPES_REQUIRE(x && y, "null vector");
    CsrF32 A{n_rows, n_cols, nnz, rowptr, col, val};
    launch_spmv_f32(A, x, y, csr_group_for(n_rows, nnz), as_stream(stream));
  });
}

int impes_spmv_st7_f64(int nx, int ny, int nz, const double* d, const double* ox, const double* oy,
                       const double* oz, const double* x, double* y, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad grid");
    IMPES_REQUIRE(d && ox && oy && oz && x && y, "null pointer");
    Stencil7<double> S{nx, ny, nz, d, ox, oy, oz};
    launch_st7_spmv_f64(S, x, y, as_stream(stream));
  });
}

int impes_l1jacobi_csr_f64(int n, int nnz, const int* rowptr, const int* col, const double* val,
                           const double* inv_relax, double omega, const double* x_in, const double* b,
                           double* x_out, void* stream) {
  return guarded(__func__, [&] {
    check_csr(n, nnz, rowptr, col, val);
    IMPES_REQUIRE(inv_relax && x_in && b && x_out, "null vector");
    IMPES_REQUIRE(x_in != x_out, "x_in and x_out must differ (Jacobi sweep)");
    CsrF64 A{n, n, nnz, rowptr, col, val};
    launch_jacobi_f64(A, x_in, b, inv_relax, omega, x_out, csr_group_for(n, nnz), as_stream(stream));
  });
}

int impes_l1jacobi_csr_f32(int n, int nnz, const int* rowptr, const int* col, const float* val,
                           const float* inv_relax, float omega, const float* x_in, const float* b, float* x_out,
                           void* stream) {
  return guarded(__func__, [&] {
    check_csr(n, nnz, rowptr, col, val);
    IMPES_REQUIRE(inv_relax && x_in && b && x_out, "null vector");
    IMPES_REQUIRE(x_in != x_out, "x_in and x_out must differ (Jacobi sweep)");
    CsrF32 A{n, n, nnz, rowptr, col, val};
    launch_jacobi_f32(A, x_in, b, inv_relax, omega, x_out, csr_group_for(n, nnz), as_stream(stream));
  });
}

int impes_dot_f64(long long n, const double* x, const double* y, double* out, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(n >= 0 && x && y && out, "bad argument");
    DevBuf<double> work;
    work.alloc(kRedBlocks);
    launch_dot_f64(x, y, n, out, work.p, as_stream(stream));
    IMPES_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  });
}

void impes_amg_default_params(impes_amg_params* p) {
  if (!p) return;
  AmgParams d;
  p->theta = d.theta;
  p->max_levels = d.max_levels;
  p->coarsest_size = d.coarsest_size;
  p->pre_smooth = d.pre_smooth;
  p->post_smooth = d.post_smooth;
  p->omega = d.omega;
  p->equil_threshold = d.equil_threshold;
  p->dense_max = d.dense_max;
  p->coarse_sweeps = d.coarse_sweeps;
  p->anchor = d.anchor;
}

int impes_amg_create(const impes_amg_params* params, impes_amg** out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(out, "null out");
    AmgParams p;
    if (params) {
      p.theta = params->theta;
      p.max_levels = params->max_levels;
      p.coarsest_size = params->coarsest_size;
      p.pre_smooth = params->pre_smooth;
      p.post_smooth = params->post_smooth;
      p.omega = params->omega;
      p.equil_threshold = params->equil_threshold;
      p.dense_max = params->dense_max;
      p.coarse_sweeps = params->coarse_sweeps;
      p.anchor = params->anchor;
    }
    IMPES_REQUIRE(p.max_levels >= 1 && p.max_levels <= 64, "max_levels out of range");
    IMPES_REQUIRE(p.pre_smooth >= 0 && p.post_smooth >= 0, "negative sweep count");
    IMPES_REQUIRE(p.dense_max >= 1 && p.dense_max <= 512, "dense_max must be in [1, 512]");
    *out = new impes_amg(p);
  });
}

int impes_amg_destroy(impes_amg* amg) {
  return guarded(__func__, [&] { delete amg; });
}

int impes_amg_setup_csr(impes_amg* amg, int n, int nnz, const int* rowptr, const int* col, const double* val,
                        void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg, "null handle");
    check_csr(n, nnz, rowptr, col, val);
    amg->h.setup_csr(n, nnz, rowptr, col, val, as_stream(stream));
  });
}

int impes_amg_setup_st7(impes_amg* amg, int nx, int ny, int nz, const double* d, const double* ox,
                        const double* oy, const double* oz, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && d && ox && oy && oz, "null pointer");
    IMPES_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad grid");
    amg->h.setup_stencil(nx, ny, nz, d, ox, oy, oz, as_stream(stream));
  });
}

int impes_amg_update_csr(impes_amg* amg, const double* val, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && val, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "update before setup");
    IMPES_REQUIRE(!amg->h.is_stencil(), "hierarchy was set up from a stencil: use impes_amg_update_st7");
    amg->h.update_csr_values(val, as_stream(stream));
  });
}

int impes_amg_update_st7(impes_amg* amg, const double* d, const double* ox, const double* oy, const double* oz,
                         void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && d && ox && oy && oz, "null pointer");
    amg->h.update_stencil(d, ox, oy, oz, as_stream(stream));
  });
}

int impes_amg_num_levels(const impes_amg* amg, int* out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && out, "null pointer");
    *out = amg->h.num_levels();
  });
}

int impes_amg_level_info(const impes_amg* amg, int level, int* n, int* nnz, int* p_nnz, int* is_coarsest,
                         int* dense) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg, "null handle");
    IMPES_REQUIRE(level >= 0 && level < amg->h.num_levels(), "level out of range");
    LevelInfo li = amg->h.info(level);
    if (n) *n = li.n;
    if (nnz) *nnz = li.nnz;
    if (p_nnz) *p_nnz = li.p_nnz;
    if (is_coarsest) *is_coarsest = li.is_coarsest;
    if (dense) *dense = li.dense;
  });
}

int impes_amg_equilibrated(const impes_amg* amg, int* out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && out, "null pointer");
    *out = amg->h.equilibrated() ? 1 : 0;
  });
}

int impes_amg_setup_seconds(const impes_amg* amg, double* out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && out, "null pointer");
    *out = amg->h.last_setup_seconds;
  });
}

int impes_amg_export_level(const impes_amg* amg, int level, int* rowptr, int* col, double* val, int* p_rowptr,
                           int* p_col, double* p_val, signed char* cf, double* inv_relax, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg, "null handle");
    IMPES_REQUIRE(level >= 0 && level < amg->h.num_levels(), "level out of range");
    const Level& L = amg->h.level(level);
    hipStream_t s = as_stream(stream);
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (dst && src && bytes) IMPES_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    };
    cp(rowptr, L.rp.p, sizeof(int) * (L.n + 1));
    cp(col, L.ci.p, sizeof(int) * L.nnz);
    cp(val, L.v.p, sizeof(double) * L.nnz);
    cp(inv_relax, L.ir64.p, sizeof(double) * L.n);
    if (!L.coarsest) {
      cp(p_rowptr, L.prp.p, sizeof(int) * (L.n + 1));
      cp(p_col, L.pci.p, sizeof(int) * L.pnnz);
      cp(p_val, L.pv.p, sizeof(double) * L.pnnz);
      cp(cf, L.cf.p, L.n);
    }
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int impes_amg_export_scaling(const impes_amg* amg, double* dhalf_inv, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && dhalf_inv, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "not set up");
    IMPES_HIP_CHECK(hipMemcpyAsync(dhalf_inv, amg->h.dhalf_inv(), sizeof(double) * amg->h.n0(),
                                   hipMemcpyDeviceToDevice, as_stream(stream)));
    IMPES_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  });
}

int impes_amg_vcycle(impes_amg* amg, int precision, const double* b, double* x, int x_is_zero, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && b && x, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "not set up");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    hipStream_t s = as_stream(stream);
    const long long n = amg->h.n0();
    if (precision == 64) {
      amg->h.vcycle64(b, x, x_is_zero != 0, s);
    } else {
      DevBuf<float> b32, x32;
      b32.alloc(n);
      x32.alloc(n);
      launch_f64_to_f32(b32.p, b, n, s);
      if (!x_is_zero) launch_f64_to_f32(x32.p, x, n, s);
      amg->h.vcycle32(b32.p, x32.p, x_is_zero != 0, s);
      // widen back
      DevBuf<double> ones;
      ones.alloc(n);
      launch_fill_f64(ones.p, 1.0, n, s);
      launch_scale_f32_to_f64(x, ones.p, x32.p, n, s);
      IMPES_HIP_CHECK(hipStreamSynchronize(s));
    }
  });
}

int impes_amg_precondition(impes_amg* amg, int precision, const double* v, double* z, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && v && z, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "not set up");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    amg->h.precondition(precision, v, z, as_stream(stream));
  });
}

int impes_solver_create(impes_solver** out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(out, "null out");
    *out = new impes_solver();
  });
}

int impes_solver_destroy(impes_solver* solver) {
  return guarded(__func__, [&] { delete solver; });
}

static void fill_stats(impes_solve_stats* out, const SolveStats& st) {
  if (!out) return;
  out->iterations = st.iterations;
  out->restarts = st.restarts;
  out->converged = st.converged;
  out->precision = st.precision;
  out->rel_residual = st.rel_residual;
  out->bnorm = st.bnorm;
  out->r0_rel = st.r0_rel;
}

int impes_fgmres_csr(impes_solver* solver, impes_amg* amg, int precision, int n, int nnz, const int* rowptr,
                     const int* col, const double* val, const double* dinv, const double* b, double* x,
                     int use_x0, double rtol, int restart, int maxit, impes_solve_stats* stats, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(solver, "null solver");
    check_csr(n, nnz, rowptr, col, val);
    IMPES_REQUIRE(dinv && b && x, "null vector");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(rtol > 0.0 && maxit >= 1, "bad tolerance / iteration limit");
    Operator A;
    A.stencil = false;
    A.csr = CsrF64{n, n, nnz, rowptr, col, val};
    A.group = csr_group_for(n, nnz);
    A.dinv = dinv;
    A.n = n;
    SolveStats st;
    solver->f.solve(A, amg ? &amg->h : nullptr, precision, b, x, use_x0 != 0, rtol, restart, maxit, st,
                    as_stream(stream));
    fill_stats(stats, st);
  });
}

int impes_fgmres_st7(impes_solver* solver, impes_amg* amg, int precision, int nx, int ny, int nz,
                     const double* d, const double* ox, const double* oy, const double* oz, const double* dinv,
                     const double* b, double* x, int use_x0, double rtol, int restart, int maxit,
                     impes_solve_stats* stats, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(solver, "null solver");
    IMPES_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad grid");
    IMPES_REQUIRE(d && ox && oy && oz && dinv && b && x, "null pointer");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(rtol > 0.0 && maxit >= 1, "bad tolerance / iteration limit");
    Operator A;
    A.stencil = true;
    A.st = Stencil7<double>{nx, ny, nz, d, ox, oy, oz};
    A.dinv = dinv;
    A.n = (long long)nx * ny * nz;
    SolveStats st;
    solver->f.solve(A, amg ? &amg->h : nullptr, precision, b, x, use_x0 != 0, rtol, restart, maxit, st,
                    as_stream(stream));
    fill_stats(stats, st);
  });
}

}  // extern "C"
