// IMPES step kernels and their host orchestration.
#include "impes_model.h"
#include "model_dev.h"

#include <cstring>

#include <cmath>

// Products and sums are rounded separately (no fused multiply-add), as in the
// numpy restatement: the substep count is a ceil() of a global maximum that can
// sit exactly on an integer (a choked well filling its saturation window).
#pragma clang fp contract(off)

namespace impes {

namespace {
using namespace model_dev;
constexpr int kLedgerQ = 10;  // mass(6) + volume dV_w, dV_g, clamp_w, clamp_g

// ---------------------------------------------------------------- transmissibility
// simulation.py:52-74: T = 2 k1 k2/(k1 + k2 + 1e-15) * area / distance
__global__ void k_trans(GridDev G, const double* __restrict__ kx, const double* __restrict__ ky,
                        const double* __restrict__ kz, const double* __restrict__ pv_in, double* __restrict__ tx,
                        double* __restrict__ ty, double* __restrict__ tz, double* __restrict__ pv) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    double t = 0.0;
    // no face between two stacked reservoirs (their last and first x-layers)
    if (i < G.nx - 1 && !(G.env_nx > 0 && (i + 1) % G.env_nx == 0)) {
      const double a = kx[c], b = kx[c + G.sx];
      t = 2 * a * b / (a + b + 1e-15) * G.dy * G.dz / G.dx;
    }
    tx[c] = t;
    t = 0.0;
    if (j < G.ny - 1) {
      const double a = ky[c], b = ky[c + G.sy];
      t = 2 * a * b / (a + b + 1e-15) * G.dx * G.dz / G.dy;
    }
    ty[c] = t;
    t = 0.0;
    if (k < G.nz - 1) {
      const double a = kz[c], b = kz[c + 1];
      t = 2 * a * b / (a + b + 1e-15) * G.dx * G.dy / G.dz;
    }
    tz[c] = t;
    pv[c] = pv_in[c];
  }
}

// ---------------------------------------------------------------- pressure step
// simulation.py:871-882 + fluid properties at the previous pressure
__global__ void k_props_pressure(GridDev G, FluidParams F, PvtDev T, double dt, const double* __restrict__ p,
                                 const double* __restrict__ sw, const double* __restrict__ sg,
                                 const double* __restrict__ pv, double* __restrict__ mw, double* __restrict__ mo,
                                 double* __restrict__ mg, double* __restrict__ mt, double* __restrict__ rw,
                                 double* __restrict__ ro, double* __restrict__ rg, double* __restrict__ pc,
                                 double* __restrict__ acc) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    const double pp = p[c], s_w = sw[c], s_g = sg[c];
    double kro, krw, krg, muo, muw, mug;
    relperms(F, s_w, s_g, kro, krw, krg);
    PvtAt at{};
    if (T.n > 0) at = pvt_locate(T, pp);
    viscosities(F, T, pp, muo, muw, mug, &at);
    const double a = krw / muw, o = kro / muo, g = krg / mug;
    mw[c] = a;
    mo[c] = o;
    mg[c] = g;
    mt[c] = a + o + g;
    if (G.gravity) {
      double r1, r2, r3;
      densities(F, T, pp, r1, r2, r3, &at);
      rw[c] = r1;
      ro[c] = r2;
      rg[c] = r3;
    }
    if (F.pc_scale > 0.0) pc[c] = pc_ow(F, s_w);
    acc[c] = pv[c] * total_compressibility(F, T, pp, s_w, s_g, &at) / dt;
  }
}

// _build_pressure_matrix_vectorized (:1546-1580) + _build_pressure_rhs (:1582-1620) in face form:
// each cell visits its six faces (coefficients recomputed on both sides, no atomics).
__global__ void k_assemble(GridDev G, FluidParams F, const double* __restrict__ p, const double* __restrict__ tx,
                           const double* __restrict__ ty, const double* __restrict__ tz,
                           const double* __restrict__ mw, const double* __restrict__ mo,
                           const double* __restrict__ mg, const double* __restrict__ mt,
                           const double* __restrict__ rw, const double* __restrict__ ro,
                           const double* __restrict__ rg, const double* __restrict__ pc,
                           const double* __restrict__ acc, double* __restrict__ diag, double* __restrict__ ox,
                           double* __restrict__ oy, double* __restrict__ oz, double* __restrict__ b,
                           double* __restrict__ extra, double* __restrict__ dinv) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  const bool with_pc = F.pc_scale > 0.0;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    const double pcc = p[c];
    double d = 0.0, qg = 0.0, qpc = 0.0;
    const double* T[3] = {tx, ty, tz};
    const long long st[3] = {G.sx, G.sy, 1};
    const int idx[3] = {i, j, k};
    const int dim[3] = {G.nx, G.ny, G.nz};
    double o_plus[3] = {0.0, 0.0, 0.0};
    for (int ax = 0; ax < 3; ++ax) {
      // face to the upper neighbour: this cell is the lower one
      if (idx[ax] < dim[ax] - 1) {
        const long long nb = c + st[ax];
        const bool up_self = (pcc - p[nb]) > 0;
        const long long u = up_self ? c : nb;
        const double tl = T[ax][c] * mt[u];
        o_plus[ax] = -tl;
        d += tl;
        if (ax == 2 && G.gravity)
          qg -= T[2][c] * F.g * G.dz * (mw[u] * rw[u] + mo[u] * ro[u] + mg[u] * rg[u]);
        if (with_pc) qpc -= T[ax][c] * mo[u] * (pc[nb] - pc[c]);
      }
      // face to the lower neighbour: this cell is the upper one
      if (idx[ax] > 0) {
        const long long nb = c - st[ax];
        const bool up_nb = (p[nb] - pcc) > 0;
        const long long u = up_nb ? nb : c;
        const double tl = T[ax][nb] * mt[u];
        d += tl;
        if (ax == 2 && G.gravity)
          qg += T[2][nb] * F.g * G.dz * (mw[u] * rw[u] + mo[u] * ro[u] + mg[u] * rg[u]);
        if (with_pc) qpc += T[ax][nb] * mo[u] * (pc[c] - pc[nb]);
      }
    }
    ox[c] = o_plus[0];
    oy[c] = o_plus[1];
    oz[c] = o_plus[2];
    const double a = acc[c];
    const double dd = d + a;
    diag[c] = dd;
    dinv[c] = 1.0 / dd;
    const double e = qg + qpc;
    extra[c] = e;
    b[c] = a * pcc + e;
  }
}

// _calculate_well_terms (:1622-1660), applied in well order by one thread.
// One thread per well computes its diagonal and right-hand-side increments (into
// ws[2 w], ws[2 w + 1]) and report slots; one thread then adds the increments in
// well order (the serial loop's accumulation, bit for bit). One workgroup.
__global__ __launch_bounds__(256) void k_wells_pressure(FluidParams F, int nw, const WellDesc* __restrict__ W,
                                                        const double* __restrict__ p, const double* __restrict__ mt,
                                                        double* __restrict__ diag, double* __restrict__ b,
                                                        double* __restrict__ dinv, double* __restrict__ rep,
                                                        double* __restrict__ ws) {
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    double* R = rep + kReportHead + kWellSlots * w;
    const WellDesc& d = W[w];
    const int c = d.cell;
    double dd = 0.0, db = 0.0;
    R[kWMode] = -1.0;
    R[kWTerm] = 0.0;
    R[kWQRate] = 0.0;
    if (d.control == 0) {
      const double rate = (d.value * F.rate_scale) / kDay * (d.producer ? -1.0 : 1.0);
      bool bhp = false;
      if (d.producer && d.has_bhp_min) {
        const double denom = d.wi * mt[c] + 1e-12;
        const double preq = p[c] - rate / denom;
        const double pmin = d.bhp_min * 1e6;
        if (preq < pmin) {
          const double term = d.wi * mt[c];
          R[kWMode] = 1.0;
          R[kWTerm] = term;
          dd = term;
          db = term * pmin;
          bhp = true;
        }
      }
      if (!bhp) {
        R[kWMode] = 0.0;
        R[kWQRate] = rate;
        db = rate;
      }
    } else if (d.control == 1) {
      const double term = d.wi * mt[c];
      R[kWTerm] = term;
      dd = term;
      db = term * (d.value * 1e6);
    }
    ws[2 * w] = dd;
    ws[2 * w + 1] = db;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 0; w < nw; ++w) {
    const int c = W[w].cell;
    diag[c] += ws[2 * w];
    b[c] += ws[2 * w + 1];
  }
  for (int w = 0; w < nw; ++w) dinv[W[w].cell] = 1.0 / diag[W[w].cell];
}

// ---------------------------------------------------------------- transport
// simulation.py:973-1006: mobilities at the new pressure, densities, o-g capillarity
__global__ void k_props_sat(GridDev G, FluidParams F, PvtDev T, const double* __restrict__ p,
                            const double* __restrict__ sw, const double* __restrict__ sg, double* __restrict__ mw,
                            double* __restrict__ mo, double* __restrict__ mg, double* __restrict__ rw,
                            double* __restrict__ rg, double* __restrict__ pcog) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    const double pp = p[c];
    double kro, krw, krg, muo, muw, mug;
    relperms(F, sw[c], sg[c], kro, krw, krg);
    PvtAt at{};
    if (T.n > 0) at = pvt_locate(T, pp);
    viscosities(F, T, pp, muo, muw, mug, &at);
    mw[c] = krw / muw;
    mo[c] = kro / muo;
    mg[c] = krg / mug;
    if (G.gravity) {
      double r1, r2, r3;
      densities(F, T, pp, r1, r2, r3, &at);
      rw[c] = r1;
      rg[c] = r3;
    }
    if (F.pcog_scale > 0.0) pcog[c] = pc_og(F, sg[c]);
  }
}

// simulation.py:985-1042: upwind phase fluxes (upwind by the pressure difference only),
// gravity potential with arithmetic-mean densities, o-g capillarity on gas; divergence.
__global__ void k_flux(GridDev G, FluidParams F, const double* __restrict__ p, const double* __restrict__ tx,
                       const double* __restrict__ ty, const double* __restrict__ tz, const double* __restrict__ mw,
                       const double* __restrict__ mg, const double* __restrict__ rw, const double* __restrict__ rg,
                       const double* __restrict__ pcog, double* __restrict__ div_w, double* __restrict__ div_g) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  const bool with_pcog = F.pcog_scale > 0.0;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    const double* T[3] = {tx, ty, tz};
    const long long st[3] = {G.sx, G.sy, 1};
    const int idx[3] = {i, j, k};
    const int dim[3] = {G.nx, G.ny, G.nz};
    double dw = 0.0, dg = 0.0;
    for (int ax = 0; ax < 3; ++ax) {
      for (int side = 0; side < 2; ++side) {
        long long lo, hi;
        if (side == 0) {
          if (idx[ax] >= dim[ax] - 1) continue;
          lo = c;
          hi = c + st[ax];
        } else {
          if (idx[ax] <= 0) continue;
          lo = c - st[ax];
          hi = c;
        }
        const double dp = p[lo] - p[hi];
        const long long u = dp > 0 ? lo : hi;
        double pot_w = dp, pot_g = dp;
        if (ax == 2 && G.gravity) {
          pot_w = dp + F.g * (0.5 * (rw[lo] + rw[hi])) * G.dz;
          pot_g = dp + F.g * (0.5 * (rg[lo] + rg[hi])) * G.dz;
        }
        const double t = T[ax][lo];
        const double fw = t * mw[u] * pot_w;
        const double fg = with_pcog ? t * mg[u] * (pot_g - (pcog[hi] - pcog[lo])) : t * mg[u] * pot_g;
        if (side == 0) { dw += fw; dg += fg; }
        else { dw -= fw; dg -= fg; }
      }
    }
    div_w[c] = dw;
    div_g[c] = dg;
  }
}

struct SatWellCtx {
  GridDev G;
  FluidParams F;
  PvtDev T;
  double dt;
  int nw;
  const WellDesc* W;
  const double* p_prev;
  const double* p;
  const double* sw;
  const double* sg;
  const double* pv;
  const double* mw;
  const double* mo;
  const double* mg;
  const double* mt_pres;
  const double* ox;
  const double* oy;
  const double* oz;
  const double* acc;
  const double* extra;
};

// Sum over BHP-type wells of WI*lambda*(p_bhp - p_cell) from the pressure-equation
// row of cell c (well terms excluded).
__device__ double row_balance(const SatWellCtx& X, int c, double q_rate) {
  const GridDev& G = X.G;
  int i, j, k;
  ijk(G, c, i, j, k);
  const double pc = X.p[c];
  double f = 0.0;
  if (i < G.nx - 1) f += -X.ox[c] * (pc - X.p[c + G.sx]);
  if (i > 0) f += -X.ox[c - G.sx] * (pc - X.p[c - G.sx]);
  if (j < G.ny - 1) f += -X.oy[c] * (pc - X.p[c + G.sy]);
  if (j > 0) f += -X.oy[c - G.sy] * (pc - X.p[c - G.sy]);
  if (k < G.nz - 1) f += -X.oz[c] * (pc - X.p[c + 1]);
  if (k > 0) f += -X.oz[c - 1] * (pc - X.p[c - 1]);
  return f + X.acc[c] * (pc - X.p_prev[c]) - q_rate - X.extra[c];
}

// The per-well loop of the transport step (simulation.py:1051-1347). Each well's
// terms depend only on its cell's state and on the pressure step's report, so one
// thread per well computes them into its scratch row ws[9 w ..] (cell increments
// of qw, qg, qo; oil in/out, gas in/out; volume sums of qw, qg); then one thread
// adds the rows in well order, which gives the sums and the per-cell accumulation
// of the serial loop bit for bit. One workgroup.
enum WellScratch { kWsQW = 0, kWsQG, kWsQO, kWsOilIn, kWsOilOut, kWsGasIn, kWsGasOut, kWsSumQW, kWsSumQG,
                   kWsLen };
static_assert(kWsLen == kWellScratch, "well scratch row length");

__global__ __launch_bounds__(256) void k_wells_sat(SatWellCtx X, double* __restrict__ qw, double* __restrict__ qg,
                                                   double* __restrict__ qo, double* __restrict__ rep,
                                                   double* __restrict__ ws) {
  const FluidParams& F = X.F;
  const bool pvt = X.T.n > 0;
  const double dt = X.dt;
  for (int w = threadIdx.x; w < X.nw; w += blockDim.x) {
    double dqw = 0.0, dqg = 0.0, dqo = 0.0;
    double oil_in = 0.0, oil_out = 0.0, gas_in = 0.0, gas_out = 0.0, sum_qw = 0.0, sum_qg = 0.0;
    const WellDesc& d = X.W[w];
    double* R = rep + kReportHead + kWellSlots * w;
    R[kWHasQ] = 0.0;
    R[kWBal] = NAN;
    R[kWObs + 0] = R[kWObs + 1] = R[kWObs + 2] = 0.0;
    const int c = d.cell;
    const double P = X.p[c];
    const PvtAt atP = pvt ? pvt_locate(X.T, P) : PvtAt{};  // every lookup below is at P
    const double mwc = X.mw[c], moc = X.mo[c], mgc = X.mg[c];
    const double mtc = mwc + moc + mgc;
    const double fw = mwc / (mtc + 1e-10), fg = mgc / (mtc + 1e-10), fo = moc / (mtc + 1e-10);
    const double pvc = X.pv[c];
    const double sw = X.sw[c], sg = X.sg[c];
    if (d.control == 0) {
      const double q_base = (d.value * F.rate_scale) / kDay;
      double q;
      if (!d.producer && d.surface && pvt) q = q_base * pvt_at(X.T, d.gas ? kBg : kBw, atP);
      else q = q_base;
      q *= d.producer ? -1.0 : 1.0;
      if (!d.producer) {
        double up, dn;
        if (d.gas) { up = fmax(0.0, (1.0 - F.so_r - sw) - sg); dn = sg - F.sg_cr; }
        else { up = fmax(0.0, (1.0 - F.so_r - sg) - sw); dn = sw - F.sw_cr; }
        const double dS = (dt / mx12(pvc)) * q;
        double alpha = 1.0;
        alpha = (dS > 0.0) ? fmin(alpha, up / mx12(dS)) : fmin(alpha, dn / mx12(fabs(dS)));
        q *= fmax(0.0, fmin(1.0, alpha));
        if (d.gas) {
          dqg += q;
          sum_qg += q;
          if (pvt) {
            const double bg = pvt_at(X.T, kBg, atP), rv = pvt_at(X.T, kRv, atP);
            const double qs = fmax(q, 0.0) / mx12(bg) * kDay;
            gas_in += qs * (dt / kDay);
            oil_in += rv * qs * (dt / kDay);
          }
        } else {
          dqw += q;
          sum_qw += q;
        }
      } else {
        const bool surface = d.surface && pvt;
        if (surface || pvt) {
          const double bo = pvt_at(X.T, kBo, atP), bw = pvt_at(X.T, kBw, atP), bg = pvt_at(X.T, kBg, atP);
          const double rs = pvt_at(X.T, kRs, atP), rv = pvt_at(X.T, kRv, atP);
          if (surface) {
            if (d.surface_phase == 0) q = -(q_base / mx12(fo / mx12(bo) + fg * rv / mx12(bg)));
            else if (d.surface_phase == 1) q = -(q_base * bw) / fmax(fw, 1e-8);
            else if (d.surface_phase == 2) q = -(q_base / mx12(fg / mx12(bg) + fo * rs / mx12(bo)));
            else q = -(q_base / mx12(fo / mx12(bo) + fg * rv / mx12(bg) + fw / mx12(bw)));
          }
          const double sw_up = fmax(0.0, (1.0 - F.so_r - sg) - sw), sw_dn = fmax(0.0, sw - F.sw_cr);
          const double sg_up = fmax(0.0, (1.0 - F.so_r - sw) - sg), sg_dn = fmax(0.0, sg - F.sg_cr);
          const double dSw = (dt / mx12(pvc)) * (q * fw), dSg = (dt / mx12(pvc)) * (q * fg);
          double alpha = 1.0;
          alpha = (dSw > 0.0) ? fmin(alpha, sw_up / mx12(dSw)) : fmin(alpha, sw_dn / mx12(fabs(dSw)));
          alpha = (dSg > 0.0) ? fmin(alpha, sg_up / mx12(dSg)) : fmin(alpha, sg_dn / mx12(fabs(dSg)));
          q *= fmax(0.0, fmin(1.0, alpha));
          if (d.lim_mask) {
            const double d_o = fo / mx12(bo) + fg * rv / mx12(bg);
            const double d_w = fw / mx12(bw);
            const double d_g = fg / mx12(bg) + fo * rs / mx12(bo);
            const double d_l = d_o + d_w;
            double m = INFINITY;
            bool any = false;
            if ((d.lim_mask & 1) && d_o > 0) { m = fmin(m, (d.lim[0] / kDay) / d_o); any = true; }
            if ((d.lim_mask & 2) && d_w > 0) { m = fmin(m, (d.lim[1] / kDay) / d_w); any = true; }
            if ((d.lim_mask & 4) && d_g > 0) { m = fmin(m, (d.lim[2] / kDay) / d_g); any = true; }
            if ((d.lim_mask & 8) && d_l > 0) { m = fmin(m, (d.lim[3] / kDay) / d_l); any = true; }
            if (any && fabs(q) > m) q = -m;
          }
          const double os = fabs(q) * (fo / mx12(bo) + fg * rv / mx12(bg)) * kDay;
          const double ws = fabs(q) * (fw / mx12(bw)) * kDay;
          const double gs = fabs(q) * (fg / mx12(bg) + fo * rs / mx12(bo)) * kDay;
          R[kWHasQ] = 1.0;
          R[kWQTotal] = q;
          R[kWSurf + 0] = os;
          R[kWSurf + 1] = ws;
          R[kWSurf + 2] = gs;
          R[kWSurf + 3] = os + ws;
          R[kWObs + 0] = os;
          R[kWObs + 1] = ws;
          R[kWObs + 2] = gs;
          const double qo_res = -q * fo, qg_res = -q * fg;
          oil_out += dt * (fmax(qo_res, 0.0) / mx12(bo) + rv * fmax(qg_res, 0.0) / mx12(bg));
          gas_out += dt * (fmax(qg_res, 0.0) / mx12(bg) + rs * fmax(qo_res, 0.0) / mx12(bo));
        } else {
          R[kWObs + 0] = fabs(q) * fo * kDay;
          R[kWObs + 1] = fabs(q) * fw * kDay;
          R[kWObs + 2] = fabs(q) * fg * kDay;
        }
        dqw += q * fw;
        dqg += q * fg;
        dqo += q * fo;
        sum_qw += q * fw;
        sum_qg += q * fg;
      }
    } else if (d.control == 1) {
      // WI*mob_t*(p_cell - p_bhp) through the row balance when this is the only
      // well with a pressure-equation diagonal term in the cell
      int nterm = 0;
      double q_rate = 0.0;
      for (int v = 0; v < X.nw; ++v) {
        if (X.W[v].cell != c) continue;
        const double* Rv = rep + kReportHead + kWellSlots * v;
        if (Rv[kWTerm] != 0.0) ++nterm;
        q_rate += Rv[kWQRate];
      }
      double q;
      if (nterm == 1) {
        const double B = row_balance(X, c, q_rate);
        R[kWBal] = B;
        q = -B * mtc / X.mt_pres[c];
      } else {
        q = d.wi * mtc * (P - d.value * 1e6);
      }
      if (!d.producer) {
        if (d.gas) {
          dqg += -q;
          sum_qg += -q;
          if (pvt) {
            const double bg = pvt_at(X.T, kBg, atP), rv = pvt_at(X.T, kRv, atP);
            const double qs = fmax(-q, 0.0) / mx12(bg) * kDay;
            gas_in += qs * (dt / kDay);
            oil_in += rv * qs * (dt / kDay);
          }
        } else {
          dqw += -q;
          sum_qw += -q;
        }
      } else {
        dqw += -q * fw;
        dqg += -q * fg;
        dqo += -q * fo;
        sum_qw += -q * fw;
        sum_qg += -q * fg;
        const double qa = fabs(q);
        if (pvt) {
          const double bo = pvt_at(X.T, kBo, atP), bg = pvt_at(X.T, kBg, atP);
          const double rs = pvt_at(X.T, kRs, atP), rv = pvt_at(X.T, kRv, atP), bw = pvt_at(X.T, kBw, atP);
          const double qo_res = qa * fo, qg_res = qa * fg;
          oil_out += dt * (qo_res / mx12(bo) + rv * qg_res / mx12(bg));
          gas_out += dt * (qg_res / mx12(bg) + rs * qo_res / mx12(bo));
          R[kWObs + 0] = qa * (fo / mx12(bo) + fg * rv / mx12(bg)) * kDay;
          R[kWObs + 1] = qa * (fw / mx12(bw)) * kDay;
          R[kWObs + 2] = qa * (fg / mx12(bg) + fo * rs / mx12(bo)) * kDay;
        } else {
          R[kWObs + 0] = qa * fo * kDay;
          R[kWObs + 1] = qa * fw * kDay;
          R[kWObs + 2] = qa * fg * kDay;
        }
      }
    }
    double* r = ws + (long long)kWsLen * w;
    r[kWsQW] = dqw;
    r[kWsQG] = dqg;
    r[kWsQO] = dqo;
    r[kWsOilIn] = oil_in;
    r[kWsOilOut] = oil_out;
    r[kWsGasIn] = gas_in;
    r[kWsGasOut] = gas_out;
    r[kWsSumQW] = sum_qw;
    r[kWsSumQG] = sum_qg;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double oil_in = 0.0, oil_out = 0.0, gas_in = 0.0, gas_out = 0.0, sum_qw = 0.0, sum_qg = 0.0;
  for (int w = 0; w < X.nw; ++w) {
    const double* r = ws + (long long)kWsLen * w;
    const int c = X.W[w].cell;
    qw[c] += r[kWsQW];
    qg[c] += r[kWsQG];
    qo[c] += r[kWsQO];
    oil_in += r[kWsOilIn];
    oil_out += r[kWsOilOut];
    gas_in += r[kWsGasIn];
    gas_out += r[kWsGasOut];
    sum_qw += r[kWsSumQW];
    sum_qg += r[kWsSumQG];
  }
  rep[kCompOIn] = oil_in;
  rep[kCompOOut] = oil_out;
  rep[kCompGIn] = gas_in;
  rep[kCompGOut] = gas_out;
  rep[kVolQW] = dt * sum_qw;
  rep[kVolQG] = dt * sum_qg;
}

// NaN/inf-propagating max (torch.max semantics)
__device__ __forceinline__ double nmax(double a, double b) {
  if (a != a || b != b) return NAN;
  return a > b ? a : b;
}

// simulation.py:1352-1373: saturation increments and the per-cell substep demand
// (also stored per cell in rat when the grid holds several reservoirs)
__global__ void k_rat(GridDev G, FluidParams F, double dt, const double* __restrict__ sw, const double* __restrict__ sg,
                      const double* __restrict__ pv, const double* __restrict__ qw, const double* __restrict__ qg,
                      const double* __restrict__ div_w, const double* __restrict__ div_g, double* __restrict__ dsw,
                      double* __restrict__ dsg, double* __restrict__ part, double* __restrict__ rat) {
  __shared__ double lds[kBlock / kWave];
  const long long n = (long long)G.nx * G.ny * G.nz;
  double m = 0.0;
  const double eps = 1e-12;
  CELL_LOOP(c, n) {
    const double k = dt / pv[c];
    const double a = k * (qw[c] - div_w[c]);
    const double g = k * (qg[c] - div_g[c]);
    dsw[c] = a;
    dsg[c] = g;
    const double s_w = sw[c], s_g = sg[c];
    const double sw_up = fmax(1.0 - F.so_r - s_g - s_w, 0.0), sw_dn = fmax(s_w - F.sw_cr, 0.0);
    const double sg_up = fmax(1.0 - F.so_r - s_w - s_g, 0.0), sg_dn = fmax(s_g - F.sg_cr, 0.0);
    const double rw = a > 0 ? a / fmax(sw_up, eps) : (-a) / fmax(sw_dn, eps);
    const double rg = g > 0 ? g / fmax(sg_up, eps) : (-g) / fmax(sg_dn, eps);
    const double rl = nmax(fabs(a) / F.max_ds, fabs(g) / F.max_ds);
    const double rc = nmax(nmax(rw, rg), rl);
    if (rat) rat[c] = rc;
    m = nmax(m, rc);
  }
  // block max with NaN propagation
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (int off = kWave / 2; off > 0; off >>= 1) m = nmax(m, __shfl_xor(m, off, kWave));
  if (lane == 0) lds[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int q = 1; q < kBlock / kWave; ++q) r = nmax(r, lds[q]);
    part[blockIdx.x] = r;
  }
}

// NaN-propagating max of the block's values (order-independent); valid on thread 0
__device__ double block_nmax(double m, double* lds) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (int off = kWave / 2; off > 0; off >>= 1) m = nmax(m, __shfl_xor(m, off, kWave));
  if (lane == 0) lds[wid] = m;
  __syncthreads();
  double r = lds[0];
  for (int q = 1; q < (int)(blockDim.x / kWave); ++q) r = nmax(r, lds[q]);
  return r;
}

__global__ void k_nsub(int nb, FluidParams F, const double* __restrict__ part, double* __restrict__ rep) {
  __shared__ double lds[kBlock / kWave];
  double m = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) m = nmax(m, part[q]);
  double r = block_nmax(m, lds);
  if (threadIdx.x != 0) return;
  if (!isfinite(r)) r = 0.0;  // nan_to_num(nan=0, posinf=0, neginf=0)
  double N = ceil(r);
  if (N < 1) N = 1;
  if (N > F.max_substeps) N = F.max_substeps;
  rep[kRatMax] = r;
  rep[kNSub] = N;
}

// Substep count of each stacked reservoir (the same rule as k_nsub over its own
// cells, which are contiguous: env e holds cells [e*cells, (e+1)*cells)). One block
// per reservoir; nsub[e] also lands in the report tail.
__global__ void k_nsub_env(long long cells, FluidParams F, const double* __restrict__ rat, double* __restrict__ nsub,
                           double* __restrict__ rep_tail) {
  __shared__ double lds[kBlock / kWave];
  const long long c0 = (long long)blockIdx.x * cells;
  double m = 0.0;
  for (long long c = threadIdx.x; c < cells; c += blockDim.x) m = nmax(m, rat[c0 + c]);
  double r = block_nmax(m, lds);
  if (threadIdx.x != 0) return;
  if (!isfinite(r)) r = 0.0;
  double N = ceil(r);
  if (N < 1) N = 1;
  if (N > F.max_substeps) N = F.max_substeps;
  nsub[blockIdx.x] = N;
  rep_tail[blockIdx.x] = N;
}

// one substep of simulation.py:1385-1394 (shared by the field kernel and the ledger kernel)
__device__ __forceinline__ void substep(const FluidParams& F, double dsw_sub, double dsg_sub, double& s_w, double& s_g,
                                        double& cw, double& cg) {
  const double wu = s_w + dsw_sub;
  const double gu = s_g + dsg_sub;
  const double w = fmin(fmax(wu, F.sw_cr), 1.0 - F.so_r - gu);
  const double g = fmin(fmax(gu, F.sg_cr), 1.0 - F.so_r - w);
  cw = w - wu;
  cg = g - gu;
  s_w = w;
  s_g = g;
}

// substeps, write-back and the ledgers (simulation.py:1385-1525): mass ledger at
// the new pressure and the volume ledger sum PV dS vs dt sum q + clamp corrections
__global__ void k_substeps(GridDev G, FluidParams F, PvtDev T, const double* __restrict__ rep,
                           const double* __restrict__ p, const double* __restrict__ sw, const double* __restrict__ sg,
                           const double* __restrict__ pv, const double* __restrict__ dsw,
                           const double* __restrict__ dsg, double* __restrict__ sw_out, double* __restrict__ sg_out,
                           double* __restrict__ so_out, double* __restrict__ part, const double* __restrict__ nsub,
                           long long env_cells) {
  __shared__ double lds[kBlock / kWave];
  const long long n = (long long)G.nx * G.ny * G.nz;
  const int N_all = (int)rep[kNSub];
  double acc[kLedgerQ];
#pragma unroll
  for (int q = 0; q < kLedgerQ; ++q) acc[q] = 0.0;
  CELL_LOOP(c, n) {
    const int N = nsub ? (int)nsub[c / env_cells] : N_all;  // this cell's reservoir
    const double w0 = sw[c], g0 = sg[c];
    const double a = dsw[c] / N, b = dsg[c] / N;
    double w = w0, g = g0, cw = 0.0, cg = 0.0;
    const double v = pv[c];
    for (int s = 0; s < N; ++s) {
      double c1, c2;
      substep(F, a, b, w, g, c1, c2);
      cw += v * c1;
      cg += v * c2;
    }
    sw_out[c] = w;
    sg_out[c] = g;
    if (so_out) so_out[c] = 1.0 - w - g;
    double rw, ro, rg;
    densities(F, T, p[c], rw, ro, rg);
    const double dw = w - w0, dg = g - g0;
    const double mw = dw * v * rw, mg = dg * v * rg, mo = (-(dw + dg)) * v * ro;
    acc[0] += fmax(mw, 0.0);
    acc[1] += -fmin(mw, 0.0);
    acc[2] += fmax(mo, 0.0);
    acc[3] += -fmin(mo, 0.0);
    acc[4] += fmax(mg, 0.0);
    acc[5] += -fmin(mg, 0.0);
    acc[6] += v * dw;
    acc[7] += v * dg;
    acc[8] += cw;
    acc[9] += cg;
  }
  for (int q = 0; q < kLedgerQ; ++q) {
    const double s = block_sum(acc[q], lds);
    if (threadIdx.x == 0) part[(long long)q * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_ledger_fold(int nb, const double* __restrict__ part, double* __restrict__ rep) {
  __shared__ double lds[kBlock / kWave];
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long long)q * nb + b];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) {
    const int slot[kLedgerQ] = {kMassWIn, kMassWOut, kMassOIn, kMassOOut, kMassGIn, kMassGOut,
                                kVolDW, kVolDG, kVolCW, kVolCG};
    rep[slot[q]] = s;
  }
}

// component-balance integration over the substeps (simulation.py:1396-1503), well cells only
// Component ledger of the substeps (simulation.py:1396-1503). One thread per
// (well, substep) pair: it replays the cell's clamped substeps up to its own and
// adds that substep's surface in/out terms; a fixed-order block reduction keeps the
// result independent of timing.
__global__ __launch_bounds__(256) void k_substep_ledger(SatWellCtx X, const double* __restrict__ dsw,
                                                        const double* __restrict__ dsg, double* __restrict__ rep,
                                                        const double* __restrict__ nsub, long long env_cells) {
  __shared__ double lds[256 / kWave];
  const FluidParams& F = X.F;
  const int N_all = (int)rep[kNSub];
  // stacked reservoirs: each well uses its own reservoir's count, at most max_substeps
  const int Nb = nsub ? (int)F.max_substeps : N_all;
  double oil_in = 0.0, oil_out = 0.0, gas_in = 0.0, gas_out = 0.0;
  for (int t = threadIdx.x; t < X.nw * Nb; t += blockDim.x) {
    const int w = t / Nb, sub = t % Nb;
    const WellDesc& d = X.W[w];
    const int N = nsub ? (int)nsub[d.cell / env_cells] : N_all;
    if (sub >= N) continue;
    const double dt_sub = X.dt / N;
    if (d.control != 0 && d.control != 1) continue;
    const double* R = rep + kReportHead + kWellSlots * w;
    const int c = d.cell;
    const double P = X.p[c];
    const PvtAt atP = pvt_locate(X.T, P);  // every lookup below is at P
    const double muw = pvt_at(X.T, kMuW, atP), muo = pvt_at(X.T, kMuO, atP), mug = pvt_at(X.T, kMuG, atP);
    const double bo = pvt_at(X.T, kBo, atP), bw = pvt_at(X.T, kBw, atP), bg = pvt_at(X.T, kBg, atP);
    const double rs = pvt_at(X.T, kRs, atP), rv = pvt_at(X.T, kRv, atP);
    const double a = dsw[c] / N, b = dsg[c] / N;
    double s_w = X.sw[c], s_g = X.sg[c];
    for (int q = 0; q <= sub; ++q) {
      double c1, c2;
      substep(F, a, b, s_w, s_g, c1, c2);
    }
    const double q_base = (d.value * F.rate_scale) / kDay;
    double kro, krw, krg;
    relperms(F, s_w, s_g, kro, krw, krg);
    const double mw = krw / (muw + 1e-12), mo = kro / (muo + 1e-12), mg = krg / (mug + 1e-12);
    const double mt = mw + mo + mg;
    const double fw = mw / (mt + 1e-12), fg = mg / (mt + 1e-12), fo = mo / (mt + 1e-12);
    if (d.control == 0) {
      if (!d.producer) {
        if (d.gas) {
          const double q = d.surface ? q_base * bg : q_base;
          const double qs = fmax(q, 0.0) / mx12(bg) * kDay;
          gas_in += qs * (dt_sub / kDay);
          oil_in += rv * qs * (dt_sub / kDay);
        }
      } else {
        double q;
        if (d.surface) {
          if (d.surface_phase == 0) q = -(q_base / mx12(fo / mx12(bo) + fg * rv / mx12(bg)));
          else if (d.surface_phase == 1) q = -(q_base * bw) / mx12(fw);
          else if (d.surface_phase == 2) q = -(q_base / mx12(fg / mx12(bg) + fo * rs / mx12(bo)));
          else q = -(q_base / mx12(fo / mx12(bo) + fg * rv / mx12(bg) + fw / mx12(bw)));
        } else {
          q = -q_base;
        }
        const double qo_res = fmax(-q * fo, 0.0), qg_res = fmax(-q * fg, 0.0);
        oil_out += dt_sub * (qo_res / mx12(bo) + rv * qg_res / mx12(bg));
        gas_out += dt_sub * (qg_res / mx12(bg) + rs * qo_res / mx12(bo));
      }
    } else {
      const double bal = R[kWBal];
      const double q = (bal == bal) ? -bal * mt / X.mt_pres[c] : d.wi * mt * (P - d.value * 1e6);
      if (!d.producer) {
        if (d.gas) {
          const double qs = fmax(-q, 0.0) / mx12(bg) * kDay;
          gas_in += qs * (dt_sub / kDay);
          oil_in += rv * qs * (dt_sub / kDay);
        }
      } else {
        const double qa = fabs(q);
        const double qo_res = qa * fo, qg_res = qa * fg;
        oil_out += dt_sub * (qo_res / mx12(bo) + rv * qg_res / mx12(bg));
        gas_out += dt_sub * (qg_res / mx12(bg) + rs * qo_res / mx12(bo));
      }
    }
  }
  oil_in = block_sum(oil_in, lds);
  oil_out = block_sum(oil_out, lds);
  gas_in = block_sum(gas_in, lds);
  gas_out = block_sum(gas_out, lds);
  if (threadIdx.x == 0) {
    rep[kCompOIn] += oil_in;
    rep[kCompOOut] += oil_out;
    rep[kCompGIn] += gas_in;
    rep[kCompGOut] += gas_out;
  }
}

__global__ void k_reset_wells(int nw, const WellDesc* __restrict__ W, double* __restrict__ qw, double* __restrict__ qg,
                              double* __restrict__ qo) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int w = 0; w < nw; ++w) {
    const int c = W[w].cell;
    qw[c] = 0.0;
    qg[c] = 0.0;
    qo[c] = 0.0;
  }
}

__global__ void k_eval_props(FluidParams F, PvtDev T, long long n, const double* __restrict__ p,
                             const double* __restrict__ sw, const double* __restrict__ sg, double* __restrict__ out) {
  CELL_LOOP(c, n) {
    double kro, krw, krg, mo, mw, mg, rw, ro, rg;
    relperms(F, sw[c], sg[c], kro, krw, krg);
    viscosities(F, T, p[c], mo, mw, mg);
    densities(F, T, p[c], rw, ro, rg);
    const double v[12] = {kro, krw, krg, pc_ow(F, sw[c]), pc_og(F, sg[c]), total_compressibility(F, T, p[c], sw[c], sg[c]),
                          rw, ro, rg, mo, mw, mg};
    for (int q = 0; q < 12; ++q) out[q * n + c] = v[q];
  }
}

__global__ void k_eval_pvt(PvtDev T, int key, long long n, const double* __restrict__ p, double* __restrict__ out) {
  CELL_LOOP(c, n) out[c] = pvt_eval(T, key, p[c]);
}

}  // namespace

// ---------------------------------------------------------------- Model
Model::Model(const Grid& g, const FluidParams& f, int pvt_n, const double* pvt_host, const std::vector<WellDesc>& w)
    : grid(g), F(f) {
  IMPES_REQUIRE(g.nx >= 1 && g.ny >= 1 && g.nz >= 1, "model: bad grid");
  N_ = (long long)g.nx * g.ny * g.nz;
  IMPES_REQUIRE(N_ < (1LL << 30), "model: grid too large for int32 indexing");
  env_cells_ = N_;
  gravity_ = g.dz > 0 && g.nz > 1;  // simulation.py:999,1594
  if (pvt_n > 0) {
    IMPES_REQUIRE(pvt_n >= 2 && pvt_host, "model: PVT table needs >= 2 rows");
    std::vector<double> P(pvt_host, pvt_host + pvt_n), cols(pvt_host + pvt_n, pvt_host + 9 * pvt_n);
    for (int i = 0; i + 1 < pvt_n; ++i) IMPES_REQUIRE(P[i + 1] > P[i], "model: PVT pressures must increase");
    std::vector<double> slope(8 * (pvt_n - 1));
    for (int k = 0; k < 8; ++k)
      for (int i = 0; i + 1 < pvt_n; ++i)
        slope[k * (pvt_n - 1) + i] = (cols[k * pvt_n + i + 1] - cols[k * pvt_n + i]) / (P[i + 1] - P[i]);
    pvt_P_.alloc(pvt_n);
    pvt_col_.alloc(8 * pvt_n);
    pvt_slope_.alloc(8 * (pvt_n - 1));
    IMPES_HIP_CHECK(hipMemcpy(pvt_P_.p, P.data(), sizeof(double) * pvt_n, hipMemcpyHostToDevice));
    IMPES_HIP_CHECK(hipMemcpy(pvt_col_.p, cols.data(), sizeof(double) * 8 * pvt_n, hipMemcpyHostToDevice));
    IMPES_HIP_CHECK(hipMemcpy(pvt_slope_.p, slope.data(), sizeof(double) * slope.size(), hipMemcpyHostToDevice));
    pvt.n = pvt_n;
    pvt.P = pvt_P_.p;
    pvt.col = pvt_col_.p;
    pvt.slope = pvt_slope_.p;
  }
  const long long n = N_;
  for (auto* b : {&tx, &ty, &tz, &pv, &diag, &ox, &oy, &oz, &this->b, &dinv, &acc, &extra, &mob_t}) b->alloc(n);
  for (auto* b : {&mob_w_, &mob_o_, &mob_g_, &div_w_, &div_g_, &dsw_, &dsg_, &qw_, &qg_, &qo_}) b->alloc(n);
  if (gravity_) for (auto* b : {&rho_w_, &rho_o_, &rho_g_}) b->alloc(n);
  if (F.pc_scale > 0.0) pc_.alloc(n);
  if (F.pcog_scale > 0.0) pcog_.alloc(n);
  IMPES_HIP_CHECK(hipMemset(qw_.p, 0, n * sizeof(double)));
  IMPES_HIP_CHECK(hipMemset(qg_.p, 0, n * sizeof(double)));
  IMPES_HIP_CHECK(hipMemset(qo_.p, 0, n * sizeof(double)));
  part_.alloc((size_t)kRedBlocks * kLedgerQ);
  set_wells(w, nullptr);
  IMPES_HIP_CHECK(hipStreamSynchronize(nullptr));
}

Model::Staging::~Staging() {
  if (ev) {
    (void)hipEventSynchronize(ev);
    (void)hipEventDestroy(ev);
  }
  if (p) (void)hipHostFree(p);
}

void Model::set_wells(const std::vector<WellDesc>& w, hipStream_t s) {
  for (const auto& d : w) IMPES_REQUIRE(d.cell >= 0 && d.cell < N_, "model: well outside the grid");
  nw_ = (int)w.size();
  if (nw_ > 0) {
    wells_.reserve(nw_);
    well_sc_.reserve((size_t)kWellScratch * nw_);
    if (!stage_.ev) IMPES_HIP_CHECK(hipEventCreateWithFlags(&stage_.ev, hipEventDisableTiming));
    IMPES_HIP_CHECK(hipEventSynchronize(stage_.ev));  // the last upload has left the staging
    if (stage_.cap < (size_t)nw_) {
      if (stage_.p) IMPES_HIP_CHECK(hipHostFree(stage_.p));
      stage_.p = nullptr;
      stage_.cap = 0;
      IMPES_HIP_CHECK(hipHostMalloc((void**)&stage_.p, sizeof(WellDesc) * nw_, hipHostMallocDefault));
      stage_.cap = nw_;
    }
    std::memcpy(stage_.p, w.data(), sizeof(WellDesc) * nw_);
    IMPES_HIP_CHECK(hipMemcpyAsync(wells_.p, stage_.p, sizeof(WellDesc) * nw_, hipMemcpyHostToDevice, s));
    IMPES_HIP_CHECK(hipEventRecord(stage_.ev, s));
  }
  report_.reserve(report_len());
  launch_fill_f64(report_.p, 0.0, report_len(), s);
}

void Model::set_envs(int env_nx, hipStream_t s) {
  IMPES_REQUIRE(env_nx >= 0, "model: env_nx must be >= 0");
  if (env_nx == 0 || env_nx == grid.nx) {
    env_nx_ = 0;
    n_env_ = 1;
  } else {
    IMPES_REQUIRE(grid.nx % env_nx == 0, "model: nx must be a multiple of env_nx");
    env_nx_ = env_nx;
    n_env_ = grid.nx / env_nx;
    rat_.alloc(N_);
    nsub_.alloc(n_env_);
  }
  env_cells_ = (long long)(env_nx_ > 0 ? env_nx_ : grid.nx) * grid.ny * grid.nz;
  amg_.reset();  // the operator's pattern loses the faces between reservoirs
  report_.reserve(report_len());
  launch_fill_f64(report_.p, 0.0, report_len(), s);
}

void Model::set_rock(const double* kx, const double* ky, const double* kz, const double* pv_in, hipStream_t s) {
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  k_trans<<<grid_for(N_, kBlock, 1 << 16), kBlock, 0, s>>>(G, kx, ky, kz, pv_in, tx.p, ty.p, tz.p, pv.p);
  IMPES_LAUNCH_CHECK();
}

void Model::assemble(double dt, const double* p, const double* sw, const double* sg, hipStream_t s) {
  IMPES_REQUIRE(dt > 0.0, "assemble: dt must be positive");
  dt_ = dt;
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  const int gr = grid_for(N_, kBlock, 1 << 16);
  k_props_pressure<<<gr, kBlock, 0, s>>>(G, F, pvt, dt, p, sw, sg, pv.p, mob_w_.p, mob_o_.p, mob_g_.p, mob_t.p,
                                         rho_w_.p, rho_o_.p, rho_g_.p, pc_.p, acc.p);
  IMPES_LAUNCH_CHECK();
  k_assemble<<<gr, kBlock, 0, s>>>(G, F, p, tx.p, ty.p, tz.p, mob_w_.p, mob_o_.p, mob_g_.p, mob_t.p, rho_w_.p,
                                   rho_o_.p, rho_g_.p, pc_.p, acc.p, diag.p, ox.p, oy.p, oz.p, b.p, extra.p, dinv.p);
  IMPES_LAUNCH_CHECK();
  if (nw_ > 0) {
    k_wells_pressure<<<1, 256, 0, s>>>(F, nw_, wells_.p, p, mob_t.p, diag.p, b.p, dinv.p, report_.p, well_sc_.p);
    IMPES_LAUNCH_CHECK();
  }
}

void Model::solve(int precision, double rtol, int restart, int maxit, double* p_new, SolveStats& st, hipStream_t s) {
  if (!amg_) {
    amg_ = std::make_unique<Hierarchy>(amg_prm_);
    amg_->setup_stencil(grid.nx, grid.ny, grid.nz, diag.p, ox.p, oy.p, oz.p, s);
  } else {
    amg_->update_stencil(diag.p, ox.p, oy.p, oz.p, s);
  }
  Operator A;
  A.stencil = true;
  A.st = Stencil7<double>{grid.nx, grid.ny, grid.nz, diag.p, ox.p, oy.p, oz.p};
  A.dinv = dinv.p;
  A.n = N_;
  fgmres_.solve(A, amg_.get(), precision, b.p, p_new, true, rtol, restart, maxit, st, s);
}

bool Model::solve_cg(double tol, int maxit, double* x, int& its, hipStream_t s) {
  const long long n = N_;
  cg_r_.reserve(n); cg_z_.reserve(n); cg_p_.reserve(n); cg_ap_.reserve(n);
  cg_s_.reserve(kRedBlocks + 8);
  double* work = cg_s_.p;
  double* sc = cg_s_.p + kRedBlocks;
  Stencil7<double> S{grid.nx, grid.ny, grid.nz, diag.p, ox.p, oy.p, oz.p};
  double hbuf = 0.0;
  auto read = [&](const double* dev) {
    IMPES_HIP_CHECK(hipMemcpyAsync(&hbuf, dev, sizeof(double), hipMemcpyDeviceToHost, s));
    IMPES_HIP_CHECK(hipStreamSynchronize(s));
    return hbuf;
  };
  launch_fill_f64(x, 0.0, n, s);
  launch_copy_f64(cg_r_.p, b.p, n, s);
  launch_scale_f64(cg_z_.p, dinv.p, cg_r_.p, n, s);
  launch_copy_f64(cg_p_.p, cg_z_.p, n, s);
  launch_dot_f64(cg_r_.p, cg_z_.p, n, sc, work, s);
  double rs_old = read(sc);
  its = 0;
  if (rs_old < 1e-10) return true;
  for (int it = 0; it < maxit; ++it) {
    launch_st7_spmv_f64(S, cg_p_.p, cg_ap_.p, s);
    launch_dot_f64(cg_p_.p, cg_ap_.p, n, sc + 1, work, s);
    const double alpha = rs_old / read(sc + 1);
    launch_axpy_f64(x, alpha, cg_p_.p, n, s);
    launch_axpy_f64(cg_r_.p, -alpha, cg_ap_.p, n, s);
    launch_scale_f64(cg_z_.p, dinv.p, cg_r_.p, n, s);
    launch_dot_f64(cg_r_.p, cg_z_.p, n, sc + 2, work, s);
    const double rs_new = read(sc + 2);
    its = it + 1;
    if (std::sqrt(rs_new) < tol) return true;
    launch_xpby_f64(cg_p_.p, cg_z_.p, rs_new / rs_old, n, s);
    rs_old = rs_new;
  }
  return false;
}

void Model::saturation(double dt, const double* p_prev, const double* p_new, const double* sw, const double* sg,
                       double* sw_out, double* sg_out, double* so_out, double* host_report, hipStream_t s) {
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  const int gr = grid_for(N_, kBlock, 1 << 16);
  k_props_sat<<<gr, kBlock, 0, s>>>(G, F, pvt, p_new, sw, sg, mob_w_.p, mob_o_.p, mob_g_.p, rho_w_.p, rho_g_.p,
                                    pcog_.p);
  IMPES_LAUNCH_CHECK();
  k_flux<<<gr, kBlock, 0, s>>>(G, F, p_new, tx.p, ty.p, tz.p, mob_w_.p, mob_g_.p, rho_w_.p, rho_g_.p, pcog_.p,
                               div_w_.p, div_g_.p);
  IMPES_LAUNCH_CHECK();
  SatWellCtx X{G, F, pvt, dt, nw_, wells_.p, p_prev, p_new, sw, sg, pv.p, mob_w_.p, mob_o_.p, mob_g_.p, mob_t.p,
               ox.p, oy.p, oz.p, acc.p, extra.p};
  // head slots written below must start from zero each step
  k_wells_sat<<<1, 256, 0, s>>>(X, qw_.p, qg_.p, qo_.p, report_.p, well_sc_.p);
  IMPES_LAUNCH_CHECK();
  const bool envs = n_env_ > 1;
  k_rat<<<kRedBlocks, kBlock, 0, s>>>(G, F, dt, sw, sg, pv.p, qw_.p, qg_.p, div_w_.p, div_g_.p, dsw_.p, dsg_.p,
                                      part_.p, envs ? rat_.p : nullptr);
  IMPES_LAUNCH_CHECK();
  k_nsub<<<1, kBlock, 0, s>>>(kRedBlocks, F, part_.p, report_.p);
  IMPES_LAUNCH_CHECK();
  if (envs) {
    k_nsub_env<<<n_env_, kBlock, 0, s>>>(env_cells_, F, rat_.p, nsub_.p, report_.p + report_len() - n_env_);
    IMPES_LAUNCH_CHECK();
  }
  k_substeps<<<kRedBlocks, kBlock, 0, s>>>(G, F, pvt, report_.p, p_new, sw, sg, pv.p, dsw_.p, dsg_.p, sw_out, sg_out,
                                           so_out, part_.p, envs ? nsub_.p : nullptr, env_cells_);
  IMPES_LAUNCH_CHECK();
  k_ledger_fold<<<kLedgerQ, kBlock, 0, s>>>(kRedBlocks, part_.p, report_.p);
  IMPES_LAUNCH_CHECK();
  if (pvt.n > 0 && nw_ > 0) {
    k_substep_ledger<<<1, 256, 0, s>>>(X, dsw_.p, dsg_.p, report_.p, envs ? nsub_.p : nullptr, env_cells_);
    IMPES_LAUNCH_CHECK();
  }
  if (nw_ > 0) {
    k_reset_wells<<<1, 1, 0, s>>>(nw_, wells_.p, qw_.p, qg_.p, qo_.p);
    IMPES_LAUNCH_CHECK();
  }
  IMPES_HIP_CHECK(hipMemcpyAsync(host_report, report_.p, sizeof(double) * report_len(), hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
}

void Model::transport_forward(double dt, const double* p_prev, const double* p_new, const double* sw,
                              const double* sg, hipStream_t s) {
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  const int gr = grid_for(N_, kBlock, 1 << 16);
  k_props_sat<<<gr, kBlock, 0, s>>>(G, F, pvt, p_new, sw, sg, mob_w_.p, mob_o_.p, mob_g_.p, rho_w_.p, rho_g_.p,
                                    pcog_.p);
  IMPES_LAUNCH_CHECK();
  k_flux<<<gr, kBlock, 0, s>>>(G, F, p_new, tx.p, ty.p, tz.p, mob_w_.p, mob_g_.p, rho_w_.p, rho_g_.p, pcog_.p,
                               div_w_.p, div_g_.p);
  IMPES_LAUNCH_CHECK();
  SatWellCtx X{G, F, pvt, dt, nw_, wells_.p, p_prev, p_new, sw, sg, pv.p, mob_w_.p, mob_o_.p, mob_g_.p, mob_t.p,
               ox.p, oy.p, oz.p, acc.p, extra.p};
  k_wells_sat<<<1, 256, 0, s>>>(X, qw_.p, qg_.p, qo_.p, report_.p, well_sc_.p);
  IMPES_LAUNCH_CHECK();
  k_rat<<<kRedBlocks, kBlock, 0, s>>>(G, F, dt, sw, sg, pv.p, qw_.p, qg_.p, div_w_.p, div_g_.p, dsw_.p, dsg_.p,
                                      part_.p, nullptr);
  IMPES_LAUNCH_CHECK();
  k_nsub<<<1, kBlock, 0, s>>>(kRedBlocks, F, part_.p, report_.p);
  IMPES_LAUNCH_CHECK();
  if (nw_ > 0) {
    k_reset_wells<<<1, 1, 0, s>>>(nw_, wells_.p, qw_.p, qg_.p, qo_.p);
    IMPES_LAUNCH_CHECK();
  }
}

void Model::pressure_props(double dt, const double* p, const double* sw, const double* sg, hipStream_t s) {
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  k_props_pressure<<<grid_for(N_, kBlock, 1 << 16), kBlock, 0, s>>>(G, F, pvt, dt, p, sw, sg, pv.p, mob_w_.p,
                                                                    mob_o_.p, mob_g_.p, mob_t.p, rho_w_.p, rho_o_.p,
                                                                    rho_g_.p, pc_.p, acc.p);
  IMPES_LAUNCH_CHECK();
}

void Model::eval_props(long long n, const double* p, const double* sw, const double* sg, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_eval_props<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(F, pvt, n, p, sw, sg, out);
  IMPES_LAUNCH_CHECK();
}

void Model::eval_pvt(int key, long long n, const double* p, double* out, hipStream_t s) {
  IMPES_REQUIRE(pvt.n > 0, "no PVT table loaded");
  IMPES_REQUIRE(key >= 0 && key < 8, "bad PVT key");
  if (n <= 0) return;
  k_eval_pvt<<<grid_for(n, kBlock, 1 << 16), kBlock, 0, s>>>(pvt, key, n, p, out);
  IMPES_LAUNCH_CHECK();
}

}  // namespace impes
