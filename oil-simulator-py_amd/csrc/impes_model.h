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
  kWellSlots = 12,
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

  long long n() const { return N_; }
  int num_wells() const { return nw_; }
  int report_len() const { return kReportHead + kWellSlots * nw_ + (n_env_ > 1 ? n_env_ : 0); }
  const Hierarchy* hierarchy() const { return amg_.get(); }

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
  double dt_ = 0.0;
};

}  // namespace impes
