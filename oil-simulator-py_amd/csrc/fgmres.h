// Restarted flexible GMRES in FP64 with classical Gram-Schmidt and a second,
// always-run orthogonalisation pass (CGS2; with a preconditioner and restart <= 32
// the second pass of each basis vector is lagged into the next Arnoldi step, two
// passes over the basis per step instead of three), right-preconditioned by the
// AMG V-cycle (FP64 or FP32 hierarchy). The Krylov method runs on the row-scaled system
// D^-1 A x = D^-1 b (D = diag A): its residual is a per-cell pressure
// correction in Pa, so "relative residual" compares like with like even when
// well rows exceed interior rows by 15 orders of magnitude.
#pragma once

#include "amg.h"

namespace impes {

struct Operator {
  bool stencil = false;
  Stencil7<double> st;
  CsrF64 csr;
  int group = 1;
  const double* dinv = nullptr;  // 1/diag(A), length n
  long long n = 0;
};

struct SolveStats {
  int iterations = 0;      // total Arnoldi steps
  int restarts = 0;
  int converged = 0;
  int precision = 64;      // hierarchy precision used by the preconditioner
  double rel_residual = 0; // true ||D^-1(b-Ax)|| / ||D^-1 b|| at exit
  double bnorm = 0;        // ||D^-1 b||
  double r0_rel = 0;       // relative residual of the initial guess
  int checks = 0;          // pressure-error checks run (each one preconditioner application)
  double err_estimate = 0; // last check's ||M r||_inf / ||x||_inf (estimated relative pressure error)
};

class Fgmres {
 public:
  // Pressure-error check (0 = off). Once the relative residual reaches rtol, the
  // preconditioner applied to the scaled residual, z = M r, estimates the error
  // A^-1 (b - A x) (M approximates A^-1 D); x += z refines the answer, and the
  // solve ends when ||z||_inf <= kErrAccept * etol * ||x||_inf. Otherwise the
  // residual target is lowered by the ratio of estimate to residual and the
  // iteration continues from the corrected x. The ratios of the last two checks
  // are kept: the next solve starts from the target their larger one predicts
  // (the ratio alternates from step to step when the pressure does).
  void set_error_tol(double etol) {
    etol_ = etol;
    err_ratio_[0] = err_ratio_[1] = 0.0;
  }
  double error_tol() const { return etol_; }
  void solve(const Operator& A, Hierarchy* M, int precision, const double* b, double* x, bool use_x0, double rtol,
             int restart, int maxit, SolveStats& st, hipStream_t s);

 private:
  void ensure(long long n, int m);
  void apply(const Operator& A, const double* x, double* y, hipStream_t s);
  void residual(const Operator& A, const double* x, const double* b, double* r, hipStream_t s);
  long long n_ = 0;
  int m_ = 0;
  DevBuf<double> V_, Z_, w_, bs_, small_, work_;
  bool cgs2_ = false;           // IMPES_FGMRES_CGS2=1: three-pass CGS2 also where the lagged form applies
  double* host_ = nullptr;      // pinned scalars: [0] blocking reads, [2..3] per-iteration estimates
  double* host_dev_ = nullptr;  // device address of host_ (the estimate is stored there by the device)
  hipEvent_t est_ev_[2] = {nullptr, nullptr};
  // Convergence tests are deferred: iteration j + 1 is enqueued before the host
  // waits for the residual estimate of iteration j whenever the estimates so far
  // predict that j cannot end the cycle, so the GPU does not idle on those tests.
  // A stop after j then ignores iteration j + 1's work (it only writes Hessenberg
  // column j + 1, g[j+1..j+2] and basis vector j + 2, none of which the solution
  // for j + 1 columns reads). IMPES_FGMRES_LOOKAHEAD=1 defers every test (on the
  // 1M-cell case the ignored last iteration then costs more than the waits it
  // hides: 56.0M vs 53.9M cell-steps/s); IMPES_FGMRES_LOOKAHEAD=0 defers none.
  bool lookahead_ = false;
  bool no_defer_ = false;
  double etol_ = 0.0;
  double err_ratio_[2] = {0.0, 0.0};  // estimate / relative residual at the last two checks

 public:
  Fgmres();
  ~Fgmres();
  Fgmres(const Fgmres&) = delete;
  Fgmres& operator=(const Fgmres&) = delete;
};

}  // namespace impes
