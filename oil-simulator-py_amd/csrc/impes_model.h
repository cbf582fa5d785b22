// IMPES time step on the device: pressure assembly (face form), well terms, the
// AMG-preconditioned FGMRES pressure solve, and the explicit upwind transport
// with availability chokes, surface-rate conversions, limits, substeps and
// ledgers. Restates Simulator._impes_* of src/simulator/simulation.py:822-1660.
#pragma once

#include <memory>
#include <vector>

#include "amg.h"
#include "fgmres.h"
#include "impes_phys.h"

namespace impes {

// report layout (doubles), one device->host copy per step
enum ReportSlot {
  kRatMax = 0, kNSub = 1,
  kMassWIn = 2, kMassWOut, kMassOIn, kMassOOut, kMassGIn, kMassGOut,
  kCompOIn = 8, kCompOOut, kCompGIn, kCompGOut,
  kVolDW = 12, kVolDG, kVolQW, kVolQG, kVolCW, kVolCG,
  kReportHead = 32,
};
// per-well record inside the report, after the head
enum WellSlot {
  kWMode = 0,      // -1 untouched, 0 rate, 1 bhp (pressure step)
  kWTerm = 1,      // WI*lambda_t of the pressure equation (0 for rate wells)
  kWQRate = 2,     // rate term of the pressure equation (m3/s)
  kWHasQ = 3,      // 1 if last_q_total / surface rates were set this step
  kWQTotal = 4,
  kWSurf = 5,      // 4 surface rates: oil, water, gas, liquid (m3/day)
  kWBal = 9,       // row-balance rate of a BHP well (m3/s), NaN if the direct form is used
  kWObs = 12,      // producers: surface oil, water, gas rates (m3/day) of every rate and BHP
                   // producer (formation volume factors 1 without a PVT table); 0 for injectors
  kWellSlots = 16,
};

constexpr int kWellScratch = 9;  // per-well scratch row of the transport step's well loop

struct Grid {
  int nx, ny, nz;
  double dx, dy, dz;
};

class Model {
 public:
  Model(const Grid& g, const FluidParams& f, int pvt_n, const double* pvt_host, const std::vector<WellDesc>& wells);

  // kx/ky/kz in m^2, pore volume per cell in m^3 (reservoir.py:50)
  void set_rock(const double* kx, const double* ky, const double* kz, const double* pv, hipStream_t s);
  // The grid holds nx / env_nx independent reservoirs stacked along x (env_nx x-layers
  // each): no faces between them and a substep count per reservoir (its count in the
  // report tail). 0 or nx = one reservoir. Call before set_rock.
  void set_envs(int env_nx, hipStream_t s);
  int num_envs() const { return n_env_; }
  // new well descriptors, uploaded on stream s (through pinned staging, so the
  // caller's host array may be released on return)
  void set_wells(const std::vector<WellDesc>& wells, hipStream_t s);
  void set_amg_params(const AmgParams& p) { amg_prm_ = p; amg_.reset(); }
  void reset_hierarchy() { amg_.reset(); }
  // pressure-error check of the FGMRES solve (relative, infinity norm; 0 = off)
  void set_error_tol(double etol) { fgmres_.set_error_tol(etol); }
  double error_tol() const { return fgmres_.error_tol(); }

  // Pressure equation at the current state: face coefficients, diagonal, rhs, 1/diag.
  void assemble(double dt, const double* p, const double* sw, const double* sg, hipStream_t s);
  // AMG setup (first call) or numeric update, then FGMRES. p_new holds the initial guess on entry.
  void solve(int precision, double rtol, int restart, int maxit, double* p_new, SolveStats& st, hipStream_t s);
  // Jacobi-preconditioned CG with the reference's absolute stop sqrt(r.z) < tol from x = 0
  // (simulation.py:1662-1697); returns converged.
  bool solve_cg(double tol, int maxit, double* p_new, int& iterations, hipStream_t s);
  // Explicit transport; fills the host report (head + wells) and synchronises.
  void saturation(double dt, const double* p_prev, const double* p_new, const double* sw, const double* sg,
                  double* sw_out, double* sg_out, double* so_out, double* host_report, hipStream_t s);
  // Elementwise properties for API use: out is 12 x n (kro krw krg pc pcog ct rho_w rho_o rho_g mu_o mu_w mu_g)
  void eval_props(long long n, const double* p, const double* sw, const double* sg, double* out, hipStream_t s);
  void eval_pvt(int key, long long n, const double* p, double* out, hipStream_t s);

  // ---- reverse sweep (adjoint.hip): see the C ABI for the contracts
  // zero the accumulated transmissibility and well-control cotangents
  void adjoint_reset(hipStream_t s);
  // transport of the step assembled at (p_prev, sw, sg): g_sw0/g_sg0 = cotangents of
  // (sw, sg); adds into g_p_new and g_p_prev; g_obs_host: nw x 3 cotangents of the
  // producer surface rates
  void transport_vjp(double dt, const double* p_prev, const double* p_new, const double* sw, const double* sg,
                     const double* g_sw1, const double* g_sg1, const double* g_obs_host, double* g_p_new,
                     double* g_p_prev, double* g_sw0, double* g_sg0, hipStream_t s);
  // lambda = A^-1 g with the assembled (symmetric) operator and the hierarchy
  void solve_adjoint(int precision, double rtol, int restart, int maxit, const double* g, double* lam,
                     SolveStats& st, hipStream_t s);
  // assembly of (p_prev, sw, sg) given lambda and the solved p_new: adds into
  // g_p_prev, g_sw0, g_sg0 (after transport_vjp of the same step)
  void assemble_vjp(double dt, const double* p_prev, const double* sw, const double* sg, const double* p_new,
                    const double* lam, double* g_p_prev, double* g_sw0, double* g_sg0, hipStream_t s);
  // accumulated transmissibility cotangents -> permeability cotangents (m^2)
  void rock_vjp(const double* kx, const double* ky, const double* kz, double* g_kx, double* g_ky, double* g_kz,
                hipStream_t s);
  void adjoint_values(double* host, hipStream_t s);
  // forward pieces the reverse recomputes
  void transport_forward(double dt, const double* p_prev, const double* p_new, const double* sw, const double* sg,
                         hipStream_t s);
  void pressure_props(double dt, const double* p, const double* sw, const double* sg, hipStream_t s);

  long long n() const { return N_; }
  int num_wells() const { return nw_; }
  int report_len() const { return kReportHead + kWellSlots * nw_ + (n_env_ > 1 ? n_env_ : 0); }
  const Hierarchy* hierarchy() const { return amg_.get(); }
  Hierarchy* hierarchy_mut() { return amg_.get(); }

  Grid grid;
  FluidParams F;
  PvtDev pvt;
  DevBuf<double> tx, ty, tz, pv;                      // face form (0 on faces leaving the grid)
  DevBuf<double> diag, ox, oy, oz, b, dinv, acc, extra, mob_t;

 private:
  long long N_;
  int nw_ = 0;
  int env_nx_ = 0, n_env_ = 1;
  long long env_cells_ = 0;          // cells per stacked reservoir
  DevBuf<double> rat_, nsub_;        // per-cell substep demand, per-reservoir substep count
  bool gravity_ = false;
  DevBuf<double> pvt_P_, pvt_col_, pvt_slope_;
  DevBuf<WellDesc> wells_;
  // pinned staging of the well descriptors; `ev` marks when the last upload has
  // left it
  struct Staging {
    WellDesc* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging();
  } stage_;
  DevBuf<double> well_sc_;  // per-well terms of the transport step (kWellScratch doubles per well)
  DevBuf<double> mob_w_, mob_o_, mob_g_, rho_w_, rho_o_, rho_g_, pc_;
  DevBuf<double> pcog_, div_w_, div_g_, dsw_, dsg_, qw_, qg_, qo_;
  DevBuf<double> part_, report_;
  AmgParams amg_prm_;
  std::unique_ptr<Hierarchy> amg_;
  Fgmres fgmres_;
  DevBuf<double> cg_r_, cg_z_, cg_p_, cg_ap_, cg_s_;
  // adjoint workspace (allocated on first use)
  void ensure_adjoint();
  DevBuf<double> gtx_, gty_, gtz_, g_val_;  // accumulated over the reverse sweep
  DevBuf<double> g_div_w_, g_div_g_, g_mw_, g_mo_, g_mg_, g_rw_, g_ro_, g_rg_, g_pc_, g_mt_p_, g_acc_, g_extra_;
  DevBuf<double> g_obs_, g_ws_;
  Fgmres adj_fgmres_;
  double dt_ = 0.0;
};

}  // namespace impes
