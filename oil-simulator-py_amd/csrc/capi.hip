// extern "C" entry points: argument checking, exception -> error code mapping.
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/impes_hip.h"
#include "amg.h"
#include "amg_kernels.h"
#include "fgmres.h"
#include "impes_model.h"
#include "ktimer.h"

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

// Device workspace of impes_amg_cycles_noalloc: a file-scope array that every GPU
// the library runs on holds from module load until the process ends. It serves
// hierarchies of up to kNoAllocMaxRows rows (3 padded FP64 vectors + reduction
// partials: 24 MiB + 4 KiB). Host code never passes `scratch` itself to a kernel
// (in host code it names a host-side stand-in): it uses the device address that
// hipGetSymbolAddress returns, looked up once per device.
constexpr long long kNoAllocMaxRows = 1LL << 20;
constexpr long long kNoAllocScratch = 3 * kNoAllocMaxRows + kRedBlocks + 64;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on the first call
// made with that device current)
double* scratch_on_current_device() {
  int dev = -1;
  IMPES_HIP_CHECK(hipGetDevice(&dev));
  IMPES_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    IMPES_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    IMPES_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

// impes_amg_params -> AmgParams, with range checks shared by every entry point
static AmgParams params_from_c(const impes_amg_params* c) {
  AmgParams p;
  if (c) {
    p.theta = c->theta;
    p.max_levels = c->max_levels;
    p.coarsest_size = c->coarsest_size;
    p.pre_smooth = c->pre_smooth;
    p.post_smooth = c->post_smooth;
    p.omega = c->omega;
    p.equil_threshold = c->equil_threshold;
    p.dense_max = c->dense_max;
    p.coarse_sweeps = c->coarse_sweeps;
    p.anchor = c->anchor;
    p.equil_mode = c->equil_mode;
    p.coarse_pin = c->coarse_pin;
    p.trunc = c->trunc;
    p.pmax = c->pmax;
    p.coarsening = c->coarsening;
    p.dominant = c->dominant;
  }
  IMPES_REQUIRE(p.max_levels >= 1 && p.max_levels <= 64, "max_levels out of range");
  IMPES_REQUIRE(p.pre_smooth >= 0 && p.post_smooth >= 0, "negative sweep count");
  IMPES_REQUIRE(p.dense_max >= 1 && p.dense_max <= kDenseMfmaMax, "dense_max must be in [1, 4096]");
  IMPES_REQUIRE(p.equil_mode == kEquilSymmetric || p.equil_mode == kEquilScalar, "equil_mode must be 0 or 1");
  IMPES_REQUIRE(p.coarse_pin >= kPinNever && p.coarse_pin <= kPinAuto, "coarse_pin must be 0, 1 or 2");
  IMPES_REQUIRE(p.trunc >= 0.0 && p.trunc < 1.0, "trunc must be in [0, 1)");
  IMPES_REQUIRE(p.pmax >= 0, "pmax must be >= 0");
  IMPES_REQUIRE(p.coarsening == kCoarsenSinglePass || p.coarsening == kCoarsenIterated, "coarsening must be 0 or 1");
  IMPES_REQUIRE(p.dominant == 0.0 || p.dominant >= 1.0, "dominant must be 0 or >= 1");
  return p;
}

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

int impes_csr_diagonal(int n, const int* rowptr, const int* col, const double* val, double* diag, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(n >= 0 && rowptr && col && val && diag, "null pointer");
    launch_diag_relax(n, rowptr, col, val, diag, nullptr, nullptr, as_stream(stream));
  });
}

int impes_copy_f64(long long n, const double* x, double* y, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(n >= 0 && (n == 0 || (x && y)), "bad copy arguments");
    launch_copy_f64(y, x, n, as_stream(stream));
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
  p->equil_mode = d.equil_mode;
  p->coarse_pin = d.coarse_pin;
  p->trunc = d.trunc;
  p->pmax = d.pmax;
  p->coarsening = d.coarsening;
  p->dominant = d.dominant;
}

int impes_amg_create(const impes_amg_params* params, impes_amg** out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(out, "null out");
    const AmgParams p = params_from_c(params);
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

static int vcycle_entry(const char* fn, impes_amg* amg, int precision, const double* b, double* x, int x_is_zero,
                        bool transposed, void* stream) {
  return guarded(fn, [&] {
    IMPES_REQUIRE(amg && b && x, "null pointer");
    struct Restore {
      impes::Hierarchy& h;
      ~Restore() { h.set_transposed(false); }
    } restore{amg->h};
    amg->h.set_transposed(transposed);
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
      launch_f32_to_f64(x, x32.p, n, s);
      IMPES_HIP_CHECK(hipStreamSynchronize(s));  // b32 and x32 are released on return
    }
  });
}

int impes_amg_vcycle(impes_amg* amg, int precision, const double* b, double* x, int x_is_zero, void* stream) {
  return vcycle_entry(__func__, amg, precision, b, x, x_is_zero, false, stream);
}

int impes_amg_vcycle_transposed(impes_amg* amg, int precision, const double* b, double* x, int x_is_zero,
                                void* stream) {
  return vcycle_entry(__func__, amg, precision, b, x, x_is_zero, true, stream);
}

int impes_amg_set_graphs(impes_amg* amg, int enable) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg, "null pointer");
    amg->h.set_graphs(enable != 0);
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
  out->checks = st.checks;
  out->err_estimate = st.err_estimate;
}

int impes_solver_set_error_tol(impes_solver* solver, double etol) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(solver, "null solver");
    IMPES_REQUIRE(etol >= 0.0 && etol < 1.0, "error tolerance must be in [0, 1)");
    if (etol != solver->f.error_tol()) solver->f.set_error_tol(etol);
  });
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

int impes_amg_cycles(impes_amg* amg, int precision, const double* b, double* x, int use_x0, int min_cycles,
                     double tol, int max_cycles, double* rel, int* cycles_done, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && b && x, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "not set up");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(max_cycles >= 1, "max_cycles must be >= 1");
    const int c = amg->h.cycles(precision, b, x, use_x0 != 0, min_cycles, tol, max_cycles, rel, as_stream(stream));
    if (cycles_done) *cycles_done = c;
  });
}

int impes_amg_cycles_noalloc(impes_amg* amg, int precision, const double* b, double* x, int use_x0,
                             int min_cycles, double tol, int max_cycles, double* rel, int* cycles_done,
                             void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && b && x, "null pointer");
    IMPES_REQUIRE(amg->h.num_levels() > 0, "not set up");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(max_cycles >= 1, "max_cycles must be >= 1");
    const long long n = amg->h.n0();
    IMPES_REQUIRE(n <= kNoAllocMaxRows && Hierarchy::cycle_work_doubles(n) <= kNoAllocScratch,
                  "impes_amg_cycles_noalloc: the hierarchy has " + std::to_string(n) +
                      " rows, above the static workspace's maximum of " + std::to_string(kNoAllocMaxRows) +
                      " (use impes_amg_cycles)");
    double* ws = scratch_on_current_device();
    const int c =
        amg->h.cycles_in(ws, precision, b, x, use_x0 != 0, min_cycles, tol, max_cycles, rel, as_stream(stream));
    if (cycles_done) *cycles_done = c;
  });
}

// ------------------------------------------------------------------ model
static_assert(sizeof(impes_fluid) == sizeof(FluidParams), "impes_fluid layout");
static_assert(offsetof(impes_fluid, rate_scale) == offsetof(FluidParams, rate_scale), "impes_fluid layout");
static_assert(offsetof(impes_fluid, max_substeps) == offsetof(FluidParams, max_substeps), "impes_fluid layout");
static_assert(sizeof(impes_well) == sizeof(WellDesc), "impes_well layout");
static_assert(offsetof(impes_well, lim) == offsetof(WellDesc, lim), "impes_well layout");
static_assert(offsetof(impes_well, value) == offsetof(WellDesc, value), "impes_well layout");

struct impes_model {
  Model m;
  impes_model(const Grid& g, const FluidParams& f, int pn, const double* ph, const std::vector<WellDesc>& w)
      : m(g, f, pn, ph, w) {}
};

static std::vector<WellDesc> wells_from(int nwells, const impes_well* w) {
  IMPES_REQUIRE(nwells >= 0, "negative well count");
  IMPES_REQUIRE(nwells == 0 || w, "null well array");
  std::vector<WellDesc> out(nwells);
  if (nwells) std::memcpy(out.data(), w, sizeof(WellDesc) * nwells);
  for (const auto& d : out) {
    IMPES_REQUIRE(d.producer == 0 || d.producer == 1, "well: producer flag must be 0/1");
    IMPES_REQUIRE(d.control >= 0 && d.control <= 2, "well: bad control");
    IMPES_REQUIRE(d.surface_phase >= 0 && d.surface_phase <= 3, "well: bad surface phase");
  }
  return out;
}

int impes_model_create(const impes_grid* grid, const impes_fluid* fluid, int pvt_n, const double* pvt_host,
                       int nwells, const impes_well* wells_host, impes_model** out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(grid && fluid && out, "null pointer");
    IMPES_REQUIRE(pvt_n >= 0, "negative PVT length");
    Grid g{grid->nx, grid->ny, grid->nz, grid->dx, grid->dy, grid->dz};
    FluidParams f;
    std::memcpy(&f, fluid, sizeof(f));
    IMPES_REQUIRE(f.max_substeps >= 1, "max_substeps must be >= 1");
    IMPES_REQUIRE(f.max_ds > 0.0, "max_saturation_change must be > 0");
    *out = new impes_model(g, f, pvt_n, pvt_host, wells_from(nwells, wells_host));
  });
}

int impes_model_destroy(impes_model* m) {
  return guarded(__func__, [&] { delete m; });
}

int impes_model_set_rock(impes_model* m, const double* kx, const double* ky, const double* kz, const double* pv,
                         void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && kx && ky && kz && pv, "null pointer");
    m->m.set_rock(kx, ky, kz, pv, as_stream(stream));
  });
}

int impes_model_set_envs(impes_model* m, int env_nx, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m, "null model");
    m->m.set_envs(env_nx, as_stream(stream));
  });
}

int impes_model_set_wells(impes_model* m, int nwells, const impes_well* wells_host, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m, "null model");
    m->m.set_wells(wells_from(nwells, wells_host), as_stream(stream));
  });
}

int impes_model_set_amg(impes_model* m, const impes_amg_params* params) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && params, "null pointer");
    const AmgParams p = params_from_c(params);
    m->m.set_amg_params(p);
  });
}

int impes_model_report_len(const impes_model* m, int* out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && out, "null pointer");
    *out = m->m.report_len();
  });
}

int impes_model_assemble(impes_model* m, double dt, const double* p, const double* sw, const double* sg,
                         void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p && sw && sg, "null pointer");
    m->m.assemble(dt, p, sw, sg, as_stream(stream));
  });
}

int impes_model_solve_pressure(impes_model* m, int precision, double rtol, int restart, int maxit,
                               const double* p_prev, double* p_new, impes_solve_stats* stats, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p_prev && p_new, "null pointer");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(rtol > 0.0 && maxit >= 1, "bad tolerance / iteration limit");
    hipStream_t s = as_stream(stream);
    if (p_new != p_prev)
      IMPES_HIP_CHECK(hipMemcpyAsync(p_new, p_prev, sizeof(double) * m->m.n(), hipMemcpyDeviceToDevice, s));
    SolveStats st;
    m->m.solve(precision, rtol, restart, maxit, p_new, st, s);
    fill_stats(stats, st);
  });
}

int impes_model_set_error_tol(impes_model* m, double etol) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m, "null pointer");
    IMPES_REQUIRE(etol >= 0.0 && etol < 1.0, "error tolerance must be in [0, 1)");
    if (etol != m->m.error_tol()) m->m.set_error_tol(etol);
  });
}

// ------------------------------------------------------------------ model adjoint
int impes_model_adjoint_reset(impes_model* m, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m, "null pointer");
    m->m.adjoint_reset(as_stream(stream));
  });
}

int impes_model_transport_vjp(impes_model* m, double dt, const double* p_prev, const double* p_new,
                              const double* sw, const double* sg, const double* g_sw_next, const double* g_sg_next,
                              const double* g_obs_host, double* g_p_new, double* g_p_prev, double* g_sw, double* g_sg,
                              void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p_prev && p_new && sw && sg && g_sw_next && g_sg_next && g_p_new && g_p_prev && g_sw && g_sg,
                  "null pointer");
    IMPES_REQUIRE(dt > 0.0, "dt must be positive");
    IMPES_REQUIRE(g_obs_host || m->m.num_wells() == 0, "null observation cotangents");
    m->m.transport_vjp(dt, p_prev, p_new, sw, sg, g_sw_next, g_sg_next, g_obs_host, g_p_new, g_p_prev, g_sw, g_sg,
                       as_stream(stream));
  });
}

int impes_model_solve_adjoint(impes_model* m, int precision, double rtol, int restart, int maxit, const double* g,
                              double* lam, impes_solve_stats* stats, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && g && lam, "null pointer");
    IMPES_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    IMPES_REQUIRE(rtol > 0.0 && maxit >= 1, "bad tolerance / iteration limit");
    SolveStats st;
    m->m.solve_adjoint(precision, rtol, restart, maxit, g, lam, st, as_stream(stream));
    fill_stats(stats, st);
  });
}

int impes_model_assemble_vjp(impes_model* m, double dt, const double* p_prev, const double* sw, const double* sg,
                             const double* p_new, const double* lam, double* g_p_prev, double* g_sw, double* g_sg,
                             void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p_prev && sw && sg && p_new && lam && g_p_prev && g_sw && g_sg, "null pointer");
    IMPES_REQUIRE(dt > 0.0, "dt must be positive");
    m->m.assemble_vjp(dt, p_prev, sw, sg, p_new, lam, g_p_prev, g_sw, g_sg, as_stream(stream));
  });
}

int impes_model_rock_vjp(impes_model* m, const double* kx, const double* ky, const double* kz, double* g_kx,
                         double* g_ky, double* g_kz, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && kx && ky && kz && g_kx && g_ky && g_kz, "null pointer");
    m->m.rock_vjp(kx, ky, kz, g_kx, g_ky, g_kz, as_stream(stream));
  });
}

int impes_model_adjoint_values(impes_model* m, double* g_values_host, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && (g_values_host || m->m.num_wells() == 0), "null pointer");
    m->m.adjoint_values(g_values_host, as_stream(stream));
  });
}

int impes_model_solve_cg(impes_model* m, double tol, int maxit, double* p_new, int* converged, int* iterations,
                         void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p_new, "null pointer");
    IMPES_REQUIRE(maxit >= 0, "bad iteration limit");
    int its = 0;
    const bool ok = m->m.solve_cg(tol, maxit, p_new, its, as_stream(stream));
    if (converged) *converged = ok ? 1 : 0;
    if (iterations) *iterations = its;
  });
}

int impes_model_saturation(impes_model* m, double dt, const double* p_prev, const double* p_new, const double* sw,
                           const double* sg, double* sw_out, double* sg_out, double* so_out, double* report_host,
                           void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p_prev && p_new && sw && sg && sw_out && sg_out && report_host, "null pointer");
    IMPES_REQUIRE(sw_out != sw && sg_out != sg, "saturation outputs must not alias the inputs");
    IMPES_REQUIRE(dt > 0.0, "dt must be positive");
    m->m.saturation(dt, p_prev, p_new, sw, sg, sw_out, sg_out, so_out, report_host, as_stream(stream));
  });
}

int impes_model_export(impes_model* m, int field, double* out, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && out, "null pointer");
    Model& M = m->m;
    const double* src[] = {M.diag.p, M.ox.p, M.oy.p, M.oz.p, M.b.p, M.acc.p, M.extra.p,
                           M.tx.p, M.ty.p, M.tz.p, M.pv.p, M.mob_t.p, M.dinv.p};
    IMPES_REQUIRE(field >= 0 && field < (int)(sizeof(src) / sizeof(src[0])), "bad field id");
    IMPES_HIP_CHECK(hipMemcpyAsync(out, src[field], sizeof(double) * M.n(), hipMemcpyDeviceToDevice,
                                   as_stream(stream)));
  });
}

int impes_model_eval_props(impes_model* m, long long n, const double* p, const double* sw, const double* sg,
                           double* out, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p && sw && sg && out, "null pointer");
    m->m.eval_props(n, p, sw, sg, out, as_stream(stream));
  });
}

int impes_model_eval_pvt(impes_model* m, int key, long long n, const double* p, double* out, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && p && out, "null pointer");
    m->m.eval_pvt(key, n, p, out, as_stream(stream));
  });
}

int impes_model_amg_levels(const impes_model* m, int* num_levels, int* sizes_host, int* nnz_host, int max_levels) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && num_levels, "null pointer");
    const Hierarchy* h = m->m.hierarchy();
    *num_levels = h ? h->num_levels() : 0;
    for (int l = 0; h && l < h->num_levels() && l < max_levels; ++l) {
      LevelInfo li = h->info(l);
      if (sizes_host) sizes_host[l] = li.n;
      if (nnz_host) nnz_host[l] = li.nnz;
    }
  });
}

static void level_times_into(Hierarchy& h, int precision, int cycles, double* ms_host, double* bytes_host,
                             int max_levels, double* cycle_ms, hipStream_t s) {
  IMPES_REQUIRE(ms_host && max_levels >= 0, "null pointer");
  std::vector<double> ms(h.num_levels(), 0.0);
  h.level_times(precision, cycles, ms.data(), cycle_ms, s);
  for (int l = 0; l < h.num_levels() && l < max_levels; ++l) {
    ms_host[l] = ms[l];
    if (bytes_host) bytes_host[l] = h.level_bytes(l, precision);
  }
}

int impes_model_amg_level_times(impes_model* m, int precision, int cycles, double* ms_host, double* bytes_host,
                                int max_levels, double* cycle_ms, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m, "null pointer");
    Hierarchy* h = m->m.hierarchy_mut();
    IMPES_REQUIRE(h && h->num_levels() > 0, "no pressure hierarchy yet (solve first)");
    level_times_into(*h, precision, cycles, ms_host, bytes_host, max_levels, cycle_ms, as_stream(stream));
  });
}

int impes_amg_level_times(impes_amg* amg, int precision, int cycles, double* ms_host, double* bytes_host,
                          int max_levels, double* cycle_ms, void* stream) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(amg && amg->h.num_levels() > 0, "not set up");
    level_times_into(amg->h, precision, cycles, ms_host, bytes_host, max_levels, cycle_ms, as_stream(stream));
  });
}

int impes_model_amg_setup_seconds(const impes_model* m, double* out) {
  return guarded(__func__, [&] {
    IMPES_REQUIRE(m && out, "null pointer");
    const Hierarchy* h = m->m.hierarchy();
    *out = h ? h->last_setup_seconds : 0.0;
  });
}

int impes_ktimer_config(const char* kernel, int enable) {
  return guarded(__func__, [&] { ktimer_config(kernel, enable != 0); });
}

int impes_ktimer_read(double* ms, double* bytes, long long* launches) {
  return guarded(__func__, [&] { ktimer_read(nullptr, ms, bytes, launches); });
}

}  // extern "C"
