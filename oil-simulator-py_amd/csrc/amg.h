// Ruge-Stueben AMG hierarchy living in device memory.
//
// Mirrors ClassicalAMG of src/solver/classical_amg.py:689-1392: optional
// symmetric equilibration of the fine matrix, RS coarsening with RS + F-F
// interpolation, Galerkin coarse operators, L1-Jacobi V-cycles (reference:
// V(3,2), omega = 0.7; default here V(2,2), omega = 0.8), and a dense
// DOF-0-pinned coarsest solve. Two value copies of
// every level operator are kept: FP64 and FP32 (the FP32 hierarchy is the
// preconditioner of the FP64 outer Krylov solver).
#pragma once

#include <array>
#include <vector>

#include "kernels.h"

namespace impes {

// Scaling of the fine operator when its lower-median |a_ii| is below equil_threshold.
// Symmetric: D^-1/2 A D^-1/2 with D = |diag A| (classical_amg.py:699-760), recomputed on
// every update. Scalar: A / median|a_ii|, frozen at setup; it keeps the constant vector
// as the near-nullspace that row-normalised interpolation reproduces.
constexpr int kEquilSymmetric = 0;
constexpr int kEquilScalar = 1;
// Coarsest-level DOF-0 pin: never; whenever the hierarchy has more than one level
// (classical_amg.py:1045-1069); or only when the coarsest operator is singular, i.e. every
// row has |row sum| <= 1e-8 |a_ii| (decided on the device at every setup/update).
// C/F split: one pass of strict local maxima plus lift (classical_amg.py:85-146), or
// rounds of that selection among the still-undecided points until all are decided
// (an independent-set split; coarsens far more on the dense operators of deep levels).
constexpr int kCoarsenSinglePass = 0;
constexpr int kCoarsenIterated = 1;
constexpr int kPinNever = 0;
constexpr int kPinReference = 1;
constexpr int kPinAuto = 2;

struct AmgParams {
  double theta = 0.25;
  int max_levels = 10;
  int coarsest_size = 200;
  // L1-Jacobi V-cycle (reference V(3,2), omega 0.7): V(2,2) with omega 0.8 needs as
  // many outer iterations on the IMPES matrices with one sweep fewer per level
  int pre_smooth = 2;
  int post_smooth = 2;
  double omega = 0.8;
  double equil_threshold = 1e-6;  // equilibrate when median |a_ii| is below this
  int dense_max = 500;            // largest coarsest level solved densely (<= 4096; above 160 rows
                                  // the inverse is formed on the matrix cores when the level is
                                  // SPD, else by pivoted Gauss-Jordan up to 512 rows)
  int coarse_sweeps = 50;         // Jacobi sweeps on a coarsest level above dense_max
  int anchor = -1;                // forced C point on every level (reference anchor_idx), -1 = none
  // Additions that change the preconditioner (not the converged answer).
  int equil_mode = kEquilScalar;  // scaling applied when the median test fires
  int coarse_pin = kPinAuto;      // DOF-0 pin of the coarsest level
  double trunc = 0.2;             // drop F-row weights below trunc * row max (0 = keep all)
  int pmax = 4;                   // keep at most pmax weights per F row (0 = unlimited)
  int coarsening = kCoarsenIterated;
  // rows with |a_ii| > dominant * sum_j!=i |a_ij| (BHP well cells: Dirichlet-like rows)
  // become C points, so their diagonal stays on one coarse unknown instead of being
  // spread by the row-normalised interpolation (0 = off, as in the reference)
  double dominant = 10.0;
};

struct LevelInfo {
  int n, nnz, p_nnz, is_coarsest, dense;
};

// 4-entry-padded FP32 copy of a grid transfer operator cut into wave
// tiles of whole rows; nw == 0 when a row is too long for a tile
struct Pad4 {
  int n = 0, nnz = 0, nw = 0;
  DevBuf<int> rp4, ci4, wrow, wq;
  DevBuf<float> v4;
};

struct Level {
  int n = 0, nnz = 0, group = 1;
  bool coarsest = false, dense = false;
  int pin_mode = kPinNever;  // effective pin mode of the coarsest level
  DevBuf<int> pin;           // device flag: 1 when DOF 0 is pinned
  DevBuf<int> rp, ci;
  DevBuf<double> v;
  DevBuf<float> v32;
  DevBuf<double> ir64;
  DevBuf<float> ir32;
  // FP32 operator padded to 4-entry rows (the FP32 cycle's sweeps and residuals)
  bool pad4 = false;
  int g4 = 1;
  DevBuf<int> rp4, ci4;  // ci4 released when the 16-bit offsets co4 fit
  DevBuf<short> co4;
  int nw4 = 0;               // wave tiles of the padded operator (0: none)
  DevBuf<int> wrow4, wq4;
  DevBuf<float> v4;
  DevBuf<signed char> cf;  // final C/F markers (non-coarsest levels)
  // interpolation / restriction to the next level
  int nc = 0, pnnz = 0, pgroup = 1, rgroup = 1;
  DevBuf<int> prp, pci, rrp, rci;
  DevBuf<double> pv, rv;
  DevBuf<float> pv32, rv32;
  Pad4 r4;  // FP32 restriction in the padded tiled form (read instead of rv32 when set)
  // Galerkin product data: AP = A*P (values refreshed on update), next A = R*AP
  int apnnz = 0;
  int ap_group = 1, ac_group = 1;  // lanes per output entry of the numeric products
  DevBuf<int> aprp, apci;
  DevBuf<double> apv;
  DevBuf<long long> ap_seg, ac_seg;
  DevBuf<int> ap_pe, ap_pf, ac_pe, ac_pf;
  // work vectors (x: solution, b: right-hand side, r: residual, t: Jacobi ping-pong)
  DevBuf<double> x64, b64, r64, t64;
  DevBuf<float> x32, b32, r32, t32;
  // coarsest level: [A_pinned^-1] stored in the right half of an n x 2n matrix
  DevBuf<double> aug;
  // coarsest level above 160 rows: A_pinned^-1 padded to dense_pad(n) (matrix cores)
  bool mfma = false;
  DevBuf<double> inv, gj_w, gj_p;
  DevBuf<int> gj_bad;
};

class Hierarchy {
 public:
  explicit Hierarchy(const AmgParams& p);
  ~Hierarchy();
  Hierarchy(const Hierarchy&) = delete;
  Hierarchy& operator=(const Hierarchy&) = delete;

  // Full setup from an (unscaled) CSR matrix with sorted column indices.
  void setup_csr(int n, int nnz, const int* rp, const int* ci, const double* v, hipStream_t s);
  // Setup from a 7-point stencil operator; level 0 then runs index-free.
  void setup_stencil(int nx, int ny, int nz, const double* d, const double* ox, const double* oy, const double* oz,
                     hipStream_t s);
  // New fine-level values, same pattern: re-equilibrate, re-RAP through the
  // frozen interpolation, refresh relaxation weights and the coarsest inverse
  // (classical_amg.py:1308-1392).
  void update_csr_values(const double* v, hipStream_t s);
  void update_stencil(const double* d, const double* ox, const double* oy, const double* oz, hipStream_t s);

  // One V-cycle in the equilibrated space: x <- V(b) (x_zero: start from 0).
  void vcycle64(const double* b, double* x, bool x_zero, hipStream_t s);
  void vcycle32(const float* b, float* x, bool x_zero, hipStream_t s);
  // The cycle with the pre- and post-smoothing counts swapped. For a symmetric
  // operator (R = P^T, the diagonal L1-Jacobi splitting, a symmetric coarse
  // solve) the zero-start V(nu1, nu2) cycle B satisfies B^T = V(nu2, nu1): this is
  // the transposed preconditioner (identical to the cycle when nu1 = nu2).
  void set_transposed(bool on) { transposed_ = on; }

  // Preconditioner of the row-scaled system D^-1 A x = D^-1 b:
  // z = S .* V(S .* D .* v)  with S = D^-1/2 (equilibrated) or I.
  // The V-cycle between the two scalings is replayed from a HIP graph captured on
  // first use after each setup (its buffers are fixed; updates keep them).
  void precondition(int precision, const double* v, double* z, hipStream_t s);
  // The same preconditioner in three parts, for a caller that fuses the two
  // scalings into its own kernels: write in_scale() .* v (in the hierarchy's
  // precision) to vcycle_in(precision), call vcycle_only, then z = out_scale() .*
  // vcycle_out(precision).
  const double* in_scale() const { return sin_.p; }
  const double* out_scale() const { return sout_.p; }
  void* vcycle_in(int precision);
  const void* vcycle_out(int precision) const;
  void vcycle_only(int precision, hipStream_t s);
  void set_graphs(bool on) { graphs_ = on; }
  bool graphs() const { return graphs_; }
  // Measurement: `cycles` eager V-cycles on the current V-cycle input with HIP
  // events between the phases of every level; ms[l] = the mean stream time of
  // level l's own launches per cycle (its sweeps, residual, restriction and
  // prolongation, or the coarsest solve; its children excluded), l < num_levels.
  // Then `cycles` graph replays of the same cycle: *cycle_ms = their mean.
  void level_times(int precision, int cycles, double* ms, double* cycle_ms, hipStream_t s);
  // Algorithmic HBM bytes of one visit of level l in a zero-start cycle of the given
  // precision, attributed like level_times: every operator and vector pass of the
  // level's own launches, each array once (row pointers, column indices and values
  // of A, R or P once per pass, the gathered vector once, the other vectors once).
  double level_bytes(int l, int precision) const;

  // Stationary V-cycle iteration with physical-space input/output, the semantics of
  // ClassicalAMG.solve (classical_amg.py:1108-1155) and .apply (:1157-1180):
  // bs = dh.*b, xs = x0./dh (or 0); cycle until ||bs - A0 xs||/||bs|| < tol after at
  // least min_cycles, or max_cycles; x = dh.*xs. Returns the cycles done, *rel the residual.
  int cycles(int precision, const double* b, double* x, bool use_x0, int min_cycles, double tol, int max_cycles,
             double* rel, hipStream_t s);
  // Workspace of cycles(): three n0-vectors (scaled b, scaled x, residual; each
  // padded to 32 doubles), the
  // kRedBlocks reduction partials and 4 scalars. Nothing in it needs clearing: every
  // entry is written before it is read.
  struct CycleWork {
    double *b = nullptr, *x = nullptr, *r = nullptr, *part = nullptr, *sc = nullptr;
  };
  static long long cycle_work_doubles(long long n) { return 3 * ((n + 31) / 32 * 32) + kRedBlocks + 4; }
  // The same iteration in a caller-provided workspace of cycle_work_doubles(n0)
  // doubles at `ws` (256-byte aligned): no device allocation.
  int cycles_in(double* ws, int precision, const double* b, double* x, bool use_x0, int min_cycles, double tol,
                int max_cycles, double* rel, hipStream_t s);

  int n0() const { return levels_.empty() ? 0 : levels_[0].n; }
  int num_levels() const { return (int)levels_.size(); }
  LevelInfo info(int l) const;
  bool equilibrated() const { return equil_; }
  bool coarse_pinned(hipStream_t s) const;  // reads the device pin flag
  const double* dhalf_inv() const { return dh_.p; }  // D^-1/2 or ones
  const double* diag0() const { return diag0_.p; }   // unscaled fine diagonal
  const Level& level(int l) const { return levels_[l]; }
  bool is_stencil() const { return stencil_; }
  const AmgParams& params() const { return prm_; }
  double last_setup_seconds = 0.0;

 private:
  void setup_impl(int n, int nnz, const int* rp, const int* ci, const double* v, hipStream_t s);
  void equilibrate(const double* v_unscaled, bool decide, hipStream_t s);
  // equil_ (when deciding) and dh_ from the unscaled fine diagonal
  void equil_factors(const double* diag, bool decide, hipStream_t s);
  // Galerkin products through the frozen interpolation, relaxation weights and
  // the coarsest inverse; fine_done: level 0's weights and stencil are current
  void update_levels(bool fine_done, hipStream_t s);
  void build_levels(hipStream_t s);
  void finish_level_values(int l, hipStream_t s);
  void build_coarsest(int l, hipStream_t s);
  void dense_inverse(Level& L, hipStream_t s);  // coarsest inverse on the matrix cores
  void refresh_f32(hipStream_t s);
  void build_pad4(Level& L, hipStream_t s);
  void build_pad4_transfer(int rows, int nnz, const int* rp, const int* ci, const double* v, Pad4& T, hipStream_t s);
  void refresh_stencil(hipStream_t s);
  void alloc_vectors();
  // first_sweep_done: x already holds the first pre-smoothing sweep from zero
  template <class T>
  void vcycle_level(int l, const T* b, T* x, bool x_zero, hipStream_t s, bool first_sweep_done = false);
  template <class T>
  void smooth(int l, const T* b, T*& cur, T* other, int sweeps, bool& zero, hipStream_t s);
  template <class T>
  void coarse_solve(int l, const T* b, T* x, hipStream_t s);
  void drop_graphs();
  void capture_vcycle(int precision);
  int cycles_impl(const CycleWork& w, int precision, const double* b, double* x, bool use_x0, int min_cycles,
                  double tol, int max_cycles, double* rel, hipStream_t s);

  // event marks of level_times (null outside it)
  struct Clock {
    std::vector<hipEvent_t> ev;
    std::vector<std::array<int, 3>> seg;  // level, start mark, end mark
    int used = 0;
  };
  Clock* clock_ = nullptr;
  int mark(hipStream_t s);
  void segment(int l, int m0, int m1) {
    if (clock_) clock_->seg.push_back({l, m0, m1});
  }

  bool graphs_ = true;
  bool transposed_ = false;  // eager cycles of vcycle64/32 with nu1 and nu2 swapped
  hipGraphExec_t vexec_[2] = {nullptr, nullptr};  // [0] FP32, [1] FP64 V-cycle

  AmgParams prm_;
  // device memory of builds: slab-carved small blocks, releases deferred to the
  // end of the build (declared before the buffers that may hold its blocks)
  std::shared_ptr<SetupPool> pool_ = std::make_shared<SetupPool>();
  std::vector<Level> levels_;
  bool equil_ = false;
  double scalar_ = 1.0;  // frozen scale of kEquilScalar
  DevBuf<double> v0_unscaled_;  // fine values as given
  DevBuf<double> diag0_, dh_, sin_, sout_;
  // stencil form of level 0
  bool stencil_ = false;
  int nx_ = 0, ny_ = 0, nz_ = 0;
  DevBuf<double> st_d_, st_x_, st_y_, st_z_;  // unscaled, as given
  DevBuf<double> sd_, sx_, sy_, sz_;          // equilibrated
  DevBuf<float> sd32_, sx32_, sy32_, sz32_;
  DevBuf<double> cyc_b_, cyc_x_, cyc_r_, cyc_s_;
};

// stencil helpers (amg.hip)
void launch_st7_pattern_count(int nx, int ny, int nz, int* cnt, hipStream_t s);
void launch_st7_pattern_fill(int nx, int ny, int nz, const int* rp, int* ci, hipStream_t s);
void launch_st7_to_csr_values(int nx, int ny, int nz, const double* d, const double* ox, const double* oy,
                              const double* oz, const int* rp, double* v, hipStream_t s);

}  // namespace impes
