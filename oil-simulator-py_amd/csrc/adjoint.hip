// Reverse sweep (vector-Jacobian products) of the IMPES step of impes_model.hip.
//
// One step maps (P0, S0) = (p, sw, sg) at time n to P1 = p at n+1 (pressure
// solve of A(P0, S0) P1 = b(P0, S0)) and S1 (explicit transport at P1 from S0).
// Given the cotangents of S1, of P1 (from later steps) and of the step's
// producer surface rates, the reverse step returns the cotangents of P0 and S0
// and accumulates those of the face transmissibilities (turned into per-cell
// permeability cotangents once, at the end: T depends on k only) and of the well
// controls:
//
//   transport   k_sat_vjp        clamped substeps reversed cell by cell (the
//                                clamps of the forward substeps, recomputed,
//                                are frozen); dS = dt/PV (q - div)
//               k_flux_vjp       upwind phase fluxes and their divergence, one
//                                cell gathering its six faces (no atomics)
//               k_wells_sat_vjp  per-well sources (chokes, surface conversions,
//                                limits, the BHP rate from the pressure row
//                                balance) and the observed surface rates
//               k_props_sat_vjp  mobilities, densities, o-g capillarity at P1
//   solve       A lambda = g_P1  (A symmetric: the forward hierarchy and FGMRES)
//   assembly    k_assemble_vjp   face coefficients T*lambda_t(upwind), gravity
//                                and capillary rhs terms, by face gathering
//               k_wells_p_vjp    well terms of the pressure equation
//               k_props_p_vjp    mobilities, densities, capillarity and the
//                                accumulation PV*ct/dt at P0, and b = acc*P0
//   rock        k_trans_vjp      T = 2 k1 k2/(k1 + k2 + 1e-15) area/distance
//
// Branches are those of the forward: upwind directions, substep count, clamps,
// chokes, limits and the bhp_min switch are recomputed from the same values.
// Local derivatives of the elementwise physics come from forward-mode dual
// numbers (value + K tangents) run through the same formulas as the forward
// kernels; where a formula is not differentiable (clamps, table ends) the
// one-sided choice is that of torch autograd over the same expressions
// (clamp passes the gradient at its bounds, max/min pick the selected operand).
#include <cmath>
#include <cstring>

#include "model_dev.h"

#pragma clang fp contract(off)

namespace impes {

namespace {
using namespace model_dev;

// ---------------------------------------------------------------- dual numbers
template <int K>
struct Dual {
  double v;
  double d[K];
};

template <int K>
__device__ __forceinline__ Dual<K> cst(double x) {
  Dual<K> r;
  r.v = x;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = 0.0;
  return r;
}

template <int K>
__device__ __forceinline__ Dual<K> var(double x, int k) {
  Dual<K> r = cst<K>(x);
  r.d[k] = 1.0;
  return r;
}

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v + b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b) {
  Dual<K> r = b;
  r.v = a + b.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v - b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b.v + b.d[i] * a.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] / b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a / b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -r.v * b.d[i] / b.v;
  return r;
}

// x^e; the derivative e x^(e-1) only where x carries a tangent (0 * inf stays 0)
template <int K>
__device__ __forceinline__ Dual<K> dpow(const Dual<K>& x, double e) {
  Dual<K> r;
  r.v = powe(x.v, e);  // the forward's value bit for bit
  const double c = e * powe(x.v, e - 1.0);
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = (x.d[i] == 0.0) ? 0.0 : c * x.d[i];
  return r;
}

// |x| with derivative sign(x) (0 at 0)
template <int K>
__device__ __forceinline__ Dual<K> dabs(const Dual<K>& x) {
  if (x.v > 0.0) return x;
  if (x.v < 0.0) return -x;
  return cst<K>(fabs(x.v));
}

// max(x, lo): derivative of x where x >= lo (a clamp passes the gradient at its bound)
template <int K>
__device__ __forceinline__ Dual<K> clamp_min(const Dual<K>& x, double lo) {
  return (x.v >= lo) ? x : cst<K>(lo);
}

template <int K>
__device__ __forceinline__ Dual<K> clamp01d(const Dual<K>& x) {
  if (x.v < 0.0) return cst<K>(0.0);
  if (x.v > 1.0) return cst<K>(1.0);
  return x;
}

template <int K>
__device__ __forceinline__ Dual<K> mx12d(const Dual<K>& a) {
  return clamp_min(a, 1e-12);
}

// PVT column at a located pressure; x = p / 1e6 as a dual
template <int K>
__device__ __forceinline__ Dual<K> pvt_d(const PvtDev& T, int key, const PvtAt& a, const Dual<K>& x) {
  const double* fp = T.col + key * T.n;
  if (a.side == 2) return x;
  if (a.side == -1) return cst<K>(fp[0]);
  if (a.side == 1) return cst<K>(fp[T.n - 1]);
  return T.slope[key * (T.n - 1) + a.lo] * (x - T.P[a.lo]) + fp[a.lo];
}

// ---------------------------------------------------------------- physics on duals
template <int K>
__device__ void relperms_d(const FluidParams& F, const Dual<K>& sw, const Dual<K>& sg, Dual<K>& kro, Dual<K>& krw,
                           Dual<K>& krg) {
  const double denom = 1.0 - F.sw_cr - F.so_r;
  const Dual<K> sn = clamp01d((sw - F.sw_cr) / denom);
  krw = dpow(sn, F.nw);
  if (F.stone2) {
    const Dual<K> swn = clamp01d((sw - F.sw_cr) / (denom + 1e-12));
    const double dg = 1.0 - F.sg_cr - F.so_r;
    const Dual<K> sgn = clamp01d((sg - F.sg_cr) / (dg + 1e-12));
    const Dual<K> k = F.ko_end_w * dpow(1.0 - swn, F.now_) + F.ko_end_g * dpow(1.0 - sgn, F.nog) -
                      fmin(F.ko_end_w, F.ko_end_g);
    kro = clamp_min(k, 0.0);
  } else {
    kro = dpow(1.0 - sn, F.no);
  }
  const Dual<K> sg2 = clamp01d(sg / (1.0 - F.sw_cr - F.so_r + 1e-10));
  krg = dpow(sg2, F.ng);
}

template <int K>
__device__ Dual<K> pc_ow_d(const FluidParams& F, const Dual<K>& sw) {
  const Dual<K> sn = clamp01d((sw - F.sw_cr) / (1.0 - F.sw_cr - F.so_r));
  return F.pc_scale * dpow(1.0 - sn + 1e-6, -F.pc_exp);
}

template <int K>
__device__ Dual<K> pc_og_d(const FluidParams& F, const Dual<K>& sg) {
  const double d = 1.0 - F.sw_cr - F.so_r;
  const Dual<K> sn = clamp01d(sg / (d + 1e-12));
  return F.pcog_scale * dpow(1.0 - sn + 1e-6, -F.pcog_exp);
}

// viscosities in Pa s at the located pressure a (x = p / 1e6)
template <int K>
__device__ void visc_d(const FluidParams& F, const PvtDev& T, const PvtAt& a, const Dual<K>& x, Dual<K>& mo,
                       Dual<K>& mw, Dual<K>& mg) {
  if (T.n > 0) {
    mo = pvt_d(T, kMuO, a, x) * 1e-3;
    mw = pvt_d(T, kMuW, a, x) * 1e-3;
    mg = pvt_d(T, kMuG, a, x) * 1e-3;
  } else {
    mo = cst<K>(F.mu_o);
    mw = cst<K>(F.mu_w);
    mg = cst<K>(F.mu_g);
  }
}

template <int K>
__device__ void dens_d(const FluidParams& F, const PvtDev& T, const Dual<K>& p, Dual<K>& rw, Dual<K>& ro,
                       Dual<K>& rg) {
  if (T.n == 0) {
    rw = F.rho_w_ref * (1.0 + F.c_w * (p - 1e5));
    ro = F.rho_o_ref * (1.0 + F.c_o * (p - 1e5));
    rg = F.rho_g_ref * (1.0 + F.c_g * (p - 1e5));
    return;
  }
  const PvtAt a = pvt_locate(T, p.v);
  const Dual<K> x = p / 1e6;
  const Dual<K> bw = pvt_d(T, kBw, a, x), bo = pvt_d(T, kBo, a, x), rs = pvt_d(T, kRs, a, x);
  const Dual<K> bg = pvt_d(T, kBg, a, x), rv = pvt_d(T, kRv, a, x);
  rw = F.rho_w_ref / (bw + 1e-12);
  ro = (F.rho_o_ref + rs * F.rho_g_ref) / (bo + 1e-12);
  rg = (F.rho_g_ref + rv * F.rho_o_ref) / (bg + 1e-12);
}

// total compressibility (central difference of B over +-0.1 MPa, as the forward)
template <int K>
__device__ Dual<K> ct_d(const FluidParams& F, const PvtDev& T, const Dual<K>& p, const Dual<K>& sw,
                        const Dual<K>& sg) {
  if (T.n == 0) return cst<K>(F.cf);
  const Dual<K> pm = p / 1e6;
  const Dual<K> p0 = pm * 1e6, pp = (pm + 0.1) * 1e6, pn = (pm - 0.1) * 1e6;
  const PvtAt a0 = pvt_locate(T, p0.v);
  const PvtAt ap = pvt_locate_near(T, pp.v, a0), am = pvt_locate_near(T, pn.v, a0);
  const Dual<K> x0 = p0 / 1e6, xp = pp / 1e6, xn = pn / 1e6;
  const int keys[3] = {kBw, kBo, kBg};
  Dual<K> c[3];
  for (int q = 0; q < 3; ++q) {
    const Dual<K> b0 = pvt_d(T, keys[q], a0, x0);
    const Dual<K> bp = pvt_d(T, keys[q], ap, xp);
    const Dual<K> bm = pvt_d(T, keys[q], am, xn);
    const Dual<K> dB = (bp - bm) / (2 * 0.1);
    c[q] = -(dB / (b0 + 1e-12)) / 1e6;
  }
  const Dual<K> so = 1.0 - sw - sg;
  return F.c_r + sw * c[0] + so * c[1] + sg * c[2];
}

template <int K>
__device__ __forceinline__ double dot_d(const Dual<K>& x, int k) { return x.d[k]; }

// ---------------------------------------------------------------- transport reverse
// Substeps reversed per cell. Forward (k_substeps): a = dSw/N, b = dSg/N and N times
//   wu = w + a, gu = g + b, w = min(max(wu, sw_cr), 1 - so_r - gu),
//   g = min(max(gu, sg_cr), 1 - so_r - w).
// The four clamp decisions of every substep are recorded as bits (N <= 64).
__global__ void k_sat_vjp(GridDev G, FluidParams F, double dt, const double* __restrict__ rep,
                          const double* __restrict__ sw, const double* __restrict__ sg,
                          const double* __restrict__ pv, const double* __restrict__ dsw,
                          const double* __restrict__ dsg, const double* __restrict__ g_sw1,
                          const double* __restrict__ g_sg1, double* __restrict__ g_sw0, double* __restrict__ g_sg0,
                          double* __restrict__ g_div_w, double* __restrict__ g_div_g) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  const int N = (int)rep[kNSub];
  CELL_LOOP(c, n) {
    const double a = dsw[c] / N, b = dsg[c] / N;
    double w = sw[c], g = sg[c];
    unsigned long long lw = 0, hw = 0, lg = 0, hg = 0;
    for (int s = 0; s < N; ++s) {
      const double wu = w + a;
      const double gu = g + b;
      const bool lo_w = wu < F.sw_cr;
      const double aw = lo_w ? F.sw_cr : wu;
      const double capw = 1.0 - F.so_r - gu;
      const bool hi_w = aw > capw;
      w = hi_w ? capw : aw;
      const bool lo_g = gu < F.sg_cr;
      const double ag = lo_g ? F.sg_cr : gu;
      const double capg = 1.0 - F.so_r - w;
      const bool hi_g = ag > capg;
      g = hi_g ? capg : ag;
      const unsigned long long bit = 1ull << s;
      if (lo_w) lw |= bit;
      if (hi_w) hw |= bit;
      if (lo_g) lg |= bit;
      if (hi_g) hg |= bit;
    }
    double gw = g_sw1[c], gg = g_sg1[c], ga = 0.0, gb = 0.0;
    for (int s = N - 1; s >= 0; --s) {
      const unsigned long long bit = 1ull << s;
      double ggu = 0.0, gwn = gw;
      if (hg & bit) gwn += -gg;          // g = 1 - so_r - w
      else if (!(lg & bit)) ggu += gg;   // g = gu
      double gwu = 0.0;
      if (hw & bit) ggu += -gwn;         // w = 1 - so_r - gu
      else if (!(lw & bit)) gwu += gwn;  // w = wu
      gw = gwu;
      gg = ggu;
      ga += gwu;
      gb += ggu;
    }
    g_sw0[c] = gw;
    g_sg0[c] = gg;
    const double k = dt / pv[c];
    g_div_w[c] = -(k * (ga / N));
    g_div_g[c] = -(k * (gb / N));
  }
}

// Phase fluxes (k_flux): for a face (lo, hi) with T = T[lo], dp = P1[lo] - P1[hi],
// u = dp > 0 ? lo : hi, pot = dp (+ g rho_mean dz on z faces),
// fw = T mw[u] pot_w, fg = T mg[u] (pot_g - (pcog[hi] - pcog[lo])); div[lo] += f,
// div[hi] -= f. Each cell gathers the cotangents of its own inputs over its six
// faces; the lower cell of a face adds the face's transmissibility cotangent.
__global__ void k_flux_vjp(GridDev G, FluidParams F, const double* __restrict__ p, const double* __restrict__ tx,
                           const double* __restrict__ ty, const double* __restrict__ tz,
                           const double* __restrict__ mw, const double* __restrict__ mg,
                           const double* __restrict__ rw, const double* __restrict__ rg,
                           const double* __restrict__ pcog, const double* __restrict__ g_div_w,
                           const double* __restrict__ g_div_g, double* __restrict__ g_p,
                           double* __restrict__ g_mw, double* __restrict__ g_mo, double* __restrict__ g_mg,
                           double* __restrict__ g_rw, double* __restrict__ g_rg, double* __restrict__ g_pcog,
                           double* __restrict__ gtx, double* __restrict__ gty, double* __restrict__ gtz) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  const bool with_pcog = F.pcog_scale > 0.0;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    const double* T[3] = {tx, ty, tz};
    double* GT[3] = {gtx, gty, gtz};
    const long long st[3] = {G.sx, G.sy, 1};
    const int idx[3] = {i, j, k};
    const int dim[3] = {G.nx, G.ny, G.nz};
    double gp = 0.0, gmw = 0.0, gmg = 0.0, grw = 0.0, grg = 0.0, gpc = 0.0;
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
        const double t = T[ax][lo];
        const double dp = p[lo] - p[hi];
        const long long u = dp > 0 ? lo : hi;
        double pot_w = dp, pot_g = dp;
        const bool grav = ax == 2 && G.gravity;
        if (grav) {
          pot_w = dp + F.g * (0.5 * (rw[lo] + rw[hi])) * G.dz;
          pot_g = dp + F.g * (0.5 * (rg[lo] + rg[hi])) * G.dz;
        }
        const double dpc = with_pcog ? pcog[hi] - pcog[lo] : 0.0;
        const double drive_g = with_pcog ? pot_g - dpc : pot_g;
        const double gfw = g_div_w[lo] - g_div_w[hi];
        const double gfg = g_div_g[lo] - g_div_g[hi];
        const double gpot_w = gfw * t * mw[u];
        const double gpot_g = gfg * t * mg[u];
        const double gdp = gpot_w + gpot_g;
        gp += (c == lo) ? gdp : -gdp;
        if (grav) {
          grw += gpot_w * F.g * 0.5 * G.dz;
          grg += gpot_g * F.g * 0.5 * G.dz;
        }
        if (with_pcog) gpc += (c == hi) ? -gpot_g : gpot_g;
        if (u == c) {
          gmw += gfw * t * pot_w;
          gmg += gfg * t * drive_g;
        }
        if (side == 0) GT[ax][c] += gfw * mw[u] * pot_w + gfg * mg[u] * drive_g;
      }
    }
    g_p[c] += gp;
    g_mw[c] = gmw;
    g_mo[c] = 0.0;
    g_mg[c] = gmg;
    g_rw[c] = grw;
    g_rg[c] = grg;
    g_pcog[c] = gpc;
  }
}

// Mobilities, densities and o-g capillarity of the transport step at (P1, S0):
// dual numbers over (P1, sw, sg).
__global__ void k_props_sat_vjp(GridDev G, FluidParams F, PvtDev T, const double* __restrict__ p,
                                const double* __restrict__ sw, const double* __restrict__ sg,
                                const double* __restrict__ g_mw, const double* __restrict__ g_mo,
                                const double* __restrict__ g_mg, const double* __restrict__ g_rw,
                                const double* __restrict__ g_rg, const double* __restrict__ g_pcog,
                                double* __restrict__ g_p, double* __restrict__ g_sw, double* __restrict__ g_sg) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    typedef Dual<3> D;
    const D P = var<3>(p[c], 0), W = var<3>(sw[c], 1), S = var<3>(sg[c], 2);
    D kro, krw, krg, muo, muw, mug;
    relperms_d(F, W, S, kro, krw, krg);
    const PvtAt at = T.n > 0 ? pvt_locate(T, P.v) : PvtAt{};
    visc_d(F, T, at, P / 1e6, muo, muw, mug);
    const D mw = krw / muw, mo = kro / muo, mg = krg / mug;
    double acc[3];
    for (int q = 0; q < 3; ++q) acc[q] = g_mw[c] * mw.d[q] + g_mo[c] * mo.d[q] + g_mg[c] * mg.d[q];
    if (G.gravity) {
      D rw, ro, rg;
      dens_d(F, T, P, rw, ro, rg);
      for (int q = 0; q < 3; ++q) acc[q] += g_rw[c] * rw.d[q] + g_rg[c] * rg.d[q];
    }
    if (F.pcog_scale > 0.0) {
      const D pc = pc_og_d(F, S);
      for (int q = 0; q < 3; ++q) acc[q] += g_pcog[c] * pc.d[q];
    }
    g_p[c] += acc[0];
    g_sw[c] += acc[1];
    g_sg[c] += acc[2];
  }
}

// ---------------------------------------------------------------- wells, transport
// Tangent slots of a well's inputs.
enum WellIn { kIP = 0, kIMw, kIMo, kIMg, kISw, kISg, kIVal, kIBal, kIMtP, kWellIn };
typedef Dual<kWellIn> WD;

struct WellOut {
  WD dqw, dqg;    // increments of the cell's water and gas sources (m3/s)
  WD obs[3];      // surface oil, water, gas rates of producers (m3/day)
};

// One well of k_wells_sat on duals. `bal` is the row balance (the BHP rate form
// used when row_form), mt_pres the pressure step's total mobility at the cell.
__device__ WellOut well_sat_d(const FluidParams& F, const PvtDev& T, double dt, const WellDesc& d, double P_v,
                              double mw_v, double mo_v, double mg_v, double sw_v, double sg_v, double pvc, bool row_form,
                              double bal_v, double mtp_v) {
  const bool pvt = T.n > 0;
  const WD P = var<kWellIn>(P_v, kIP);
  const WD mwc = var<kWellIn>(mw_v, kIMw), moc = var<kWellIn>(mo_v, kIMo), mgc = var<kWellIn>(mg_v, kIMg);
  const WD sw = var<kWellIn>(sw_v, kISw), sg = var<kWellIn>(sg_v, kISg);
  const WD value = var<kWellIn>(d.value, kIVal);
  const PvtAt atP = pvt ? pvt_locate(T, P_v) : PvtAt{};
  const WD x = P / 1e6;
  const WD mtc = mwc + moc + mgc;
  const WD fw = mwc / (mtc + 1e-10), fg = mgc / (mtc + 1e-10), fo = moc / (mtc + 1e-10);
  WellOut o;
  o.dqw = o.dqg = cst<kWellIn>(0.0);
  for (int q = 0; q < 3; ++q) o.obs[q] = cst<kWellIn>(0.0);
  auto surface_rates = [&](const WD& q) {
    const WD qa = dabs(q);
    if (pvt) {
      const WD bo = pvt_d(T, kBo, atP, x), bw = pvt_d(T, kBw, atP, x), bg = pvt_d(T, kBg, atP, x);
      const WD rs = pvt_d(T, kRs, atP, x), rv = pvt_d(T, kRv, atP, x);
      o.obs[0] = qa * (fo / mx12d(bo) + fg * rv / mx12d(bg)) * kDay;
      o.obs[1] = qa * (fw / mx12d(bw)) * kDay;
      o.obs[2] = qa * (fg / mx12d(bg) + fo * rs / mx12d(bo)) * kDay;
    } else {
      o.obs[0] = qa * fo * kDay;
      o.obs[1] = qa * fw * kDay;
      o.obs[2] = qa * fg * kDay;
    }
  };
  if (d.control == 0) {
    const WD q_base = (value * F.rate_scale) / kDay;
    WD q = q_base;
    if (!d.producer && d.surface && pvt) q = q_base * pvt_d(T, d.gas ? kBg : kBw, atP, x);
    q = q * (d.producer ? -1.0 : 1.0);
    if (!d.producer) {
      WD up, dn;
      if (d.gas) {
        up = clamp_min((1.0 - F.so_r - sw) - sg, 0.0);
        dn = sg - F.sg_cr;
      } else {
        up = clamp_min((1.0 - F.so_r - sg) - sw, 0.0);
        dn = sw - F.sw_cr;
      }
      const WD dS = (dt / fmax(pvc, 1e-12)) * q;
      WD alpha = cst<kWellIn>(1.0);
      const WD cand = (dS.v > 0.0) ? up / mx12d(dS) : dn / mx12d(dabs(dS));
      if (cand.v < alpha.v) alpha = cand;
      if (alpha.v < 0.0) alpha = cst<kWellIn>(0.0);
      q = q * alpha;
      if (d.gas) o.dqg = q;
      else o.dqw = q;
    } else {
      if (pvt) {
        const WD bo = pvt_d(T, kBo, atP, x), bw = pvt_d(T, kBw, atP, x), bg = pvt_d(T, kBg, atP, x);
        const WD rs = pvt_d(T, kRs, atP, x), rv = pvt_d(T, kRv, atP, x);
        if (d.surface) {
          if (d.surface_phase == 0) q = -(q_base / mx12d(fo / mx12d(bo) + fg * rv / mx12d(bg)));
          else if (d.surface_phase == 1) q = -(q_base * bw) / clamp_min(fw, 1e-8);
          else if (d.surface_phase == 2) q = -(q_base / mx12d(fg / mx12d(bg) + fo * rs / mx12d(bo)));
          else q = -(q_base / mx12d(fo / mx12d(bo) + fg * rv / mx12d(bg) + fw / mx12d(bw)));
        }
        const WD sw_up = clamp_min((1.0 - F.so_r - sg) - sw, 0.0), sw_dn = clamp_min(sw - F.sw_cr, 0.0);
        const WD sg_up = clamp_min((1.0 - F.so_r - sw) - sg, 0.0), sg_dn = clamp_min(sg - F.sg_cr, 0.0);
        const double kq = dt / fmax(pvc, 1e-12);
        const WD dSw = kq * (q * fw), dSg = kq * (q * fg);
        WD alpha = cst<kWellIn>(1.0);
        WD cand = (dSw.v > 0.0) ? sw_up / mx12d(dSw) : sw_dn / mx12d(dabs(dSw));
        if (cand.v < alpha.v) alpha = cand;
        cand = (dSg.v > 0.0) ? sg_up / mx12d(dSg) : sg_dn / mx12d(dabs(dSg));
        if (cand.v < alpha.v) alpha = cand;
        if (alpha.v < 0.0) alpha = cst<kWellIn>(0.0);
        q = q * alpha;
        if (d.lim_mask) {
          const WD d_o = fo / mx12d(bo) + fg * rv / mx12d(bg);
          const WD d_w = fw / mx12d(bw);
          const WD d_g = fg / mx12d(bg) + fo * rs / mx12d(bo);
          const WD d_l = d_o + d_w;
          const WD* dd[4] = {&d_o, &d_w, &d_g, &d_l};
          WD m = cst<kWellIn>(INFINITY);
          bool any = false;
          for (int l = 0; l < 4; ++l) {
            if (!((d.lim_mask >> l) & 1) || !(dd[l]->v > 0)) continue;
            const WD cnd = (d.lim[l] / kDay) / *dd[l];
            if (!any || cnd.v < m.v) m = cnd;
            any = true;
          }
          if (any && fabs(q.v) > m.v) q = -m;
        }
      }
      surface_rates(q);
      o.dqw = q * fw;
      o.dqg = q * fg;
    }
  } else if (d.control == 1) {
    WD q;
    if (row_form) {
      const WD bal = var<kWellIn>(bal_v, kIBal), mtp = var<kWellIn>(mtp_v, kIMtP);
      q = -bal * mtc / mtp;
    } else {
      q = d.wi * mtc * (P - value * 1e6);
    }
    if (!d.producer) {
      if (d.gas) o.dqg = -q;
      else o.dqw = -q;
    } else {
      surface_rates(q);
      o.dqw = -q * fw;
      o.dqg = -q * fg;
    }
  }
  return o;
}

// Adds g to the cotangent of the face coefficient o_f = -T_f lambda_t(u_f) of the
// pressure matrix on the face (lo, lo + st) of axis ax: T and the pressure step's
// total mobility at the upwind cell (by P0).
__device__ void face_coef_vjp(const GridDev& G, int ax, long long lo, double g_o, const double* p0,
                              const double* const* T, double* const* GT, const double* mt_p, double* g_mt_p) {
  const long long st[3] = {G.sx, G.sy, 1};
  const long long hi = lo + st[ax];
  const long long u = (p0[lo] - p0[hi]) > 0 ? lo : hi;
  const double g_tl = -g_o;
  GT[ax][lo] += g_tl * mt_p[u];
  g_mt_p[u] += g_tl * T[ax][lo];
}

struct WellVjpCtx {
  GridDev G;
  FluidParams F;
  PvtDev T;
  double dt;
  int nw;
  const WellDesc* W;
  const double* rep;
  const double* p0;
  const double* p1;
  const double* sw;
  const double* sg;
  const double* pv;
  const double* mw;  // transport mobilities (at P1)
  const double* mo;
  const double* mg;
  const double* mt_p;  // pressure-step total mobility (at P0)
  const double* ox;
  const double* oy;
  const double* oz;
  const double* acc;
  const double* tx;
  const double* ty;
  const double* tz;
  const double* g_div_w;  // -dt/PV g_dSw: the source cotangent is its negative
  const double* g_div_g;
  const double* g_obs;    // nw x 3
};

// One thread per well forms its input cotangents (dual numbers) into ws[kWellIn w
// ..]; then one thread adds them in well order to the cell arrays, including the
// row-balance chain of BHP wells:
//   bal = sum_f (-o_f)(P1_c - P1_nb) + acc_c (P1_c - P0_c) - q_rate_c - extra_c.
__global__ __launch_bounds__(256) void k_wells_sat_vjp(WellVjpCtx X, double* __restrict__ ws,
                                                       double* __restrict__ g_p1, double* __restrict__ g_p0,
                                                       double* __restrict__ g_mw, double* __restrict__ g_mo,
                                                       double* __restrict__ g_mg, double* __restrict__ g_sw,
                                                       double* __restrict__ g_sg, double* __restrict__ g_val,
                                                       double* __restrict__ g_mt_p, double* __restrict__ g_acc,
                                                       double* __restrict__ g_extra, double* __restrict__ gtx,
                                                       double* __restrict__ gty, double* __restrict__ gtz) {
  for (int w = threadIdx.x; w < X.nw; w += blockDim.x) {
    const WellDesc& d = X.W[w];
    const double* R = X.rep + kReportHead + kWellSlots * w;
    const int c = d.cell;
    const double bal = R[kWBal];
    const bool row_form = d.control == 1 && bal == bal;
    const WellOut o = well_sat_d(X.F, X.T, X.dt, d, X.p1[c], X.mw[c], X.mo[c], X.mg[c], X.sw[c], X.sg[c], X.pv[c],
                                 row_form, row_form ? bal : 0.0, X.mt_p[c]);
    const double gqw = -X.g_div_w[c], gqg = -X.g_div_g[c];
    const double* go = X.g_obs + 3 * w;
    double* r = ws + (long long)kWellIn * w;
    for (int k = 0; k < kWellIn; ++k)
      r[k] = gqw * o.dqw.d[k] + gqg * o.dqg.d[k] + go[0] * o.obs[0].d[k] + go[1] * o.obs[1].d[k] +
             go[2] * o.obs[2].d[k];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const GridDev& G = X.G;
  const double* T[3] = {X.tx, X.ty, X.tz};
  double* GT[3] = {gtx, gty, gtz};
  const double* O[3] = {X.ox, X.oy, X.oz};
  const long long st[3] = {G.sx, G.sy, 1};
  const int dim[3] = {G.nx, G.ny, G.nz};
  for (int w = 0; w < X.nw; ++w) {
    const WellDesc& d = X.W[w];
    const int c = d.cell;
    const double* r = ws + (long long)kWellIn * w;
    g_p1[c] += r[kIP];
    g_mw[c] += r[kIMw];
    g_mo[c] += r[kIMo];
    g_mg[c] += r[kIMg];
    g_sw[c] += r[kISw];
    g_sg[c] += r[kISg];
    g_val[w] += r[kIVal];
    g_mt_p[c] += r[kIMtP];
    const double gb = r[kIBal];
    if (gb == 0.0) continue;
    // row balance of cell c
    int idx[3];
    ijk(G, c, idx[0], idx[1], idx[2]);
    const double pc = X.p1[c];
    double gpc = gb * X.acc[c];
    for (int ax = 0; ax < 3; ++ax) {
      if (idx[ax] < dim[ax] - 1) {  // face (c, c + st)
        const long long nb = c + st[ax];
        const double o = O[ax][c];
        gpc += gb * (-o);
        g_p1[nb] += gb * o;
        face_coef_vjp(G, ax, c, -gb * (pc - X.p1[nb]), X.p0, T, GT, X.mt_p, g_mt_p);
      }
      if (idx[ax] > 0) {  // face (c - st, c)
        const long long nb = c - st[ax];
        const double o = O[ax][nb];
        gpc += gb * (-o);
        g_p1[nb] += gb * o;
        face_coef_vjp(G, ax, nb, -gb * (pc - X.p1[nb]), X.p0, T, GT, X.mt_p, g_mt_p);
      }
    }
    g_p1[c] += gpc;
    g_acc[c] += gb * (pc - X.p0[c]);
    g_p0[c] += -gb * X.acc[c];
    g_extra[c] += -gb;
    // q_rate of the cell: the rate terms of its unswitched rate wells
    for (int v = 0; v < X.nw; ++v) {
      const WellDesc& dv = X.W[v];
      const double* Rv = X.rep + kReportHead + kWellSlots * v;
      if (dv.cell != c || dv.control != 0 || Rv[kWMode] != 0.0) continue;
      g_val[v] += -gb * (X.F.rate_scale / kDay) * (dv.producer ? -1.0 : 1.0);
    }
  }
}

// ---------------------------------------------------------------- assembly reverse
// Face gathering of k_assemble's reverse at (P0, S0), given lambda (the adjoint
// solution: b cotangent) and P1: the matrix cotangent of the solve is
// g_A = -lambda P1^T, so a face coefficient tl = T mt[u] (diag of both cells +tl,
// off-diagonal -tl) receives g_tl = -(lambda_lo - lambda_hi)(P1_lo - P1_hi).
// The rhs cotangent is lambda plus the row-balance injection g_extra.
__global__ void k_assemble_vjp(GridDev G, FluidParams F, const double* __restrict__ p0,
                               const double* __restrict__ p1, const double* __restrict__ lam,
                               const double* __restrict__ tx, const double* __restrict__ ty,
                               const double* __restrict__ tz, const double* __restrict__ mw,
                               const double* __restrict__ mo, const double* __restrict__ mg,
                               const double* __restrict__ mt, const double* __restrict__ rw,
                               const double* __restrict__ ro, const double* __restrict__ rg,
                               const double* __restrict__ pc, const double* __restrict__ g_extra_in,
                               double* __restrict__ g_mt, double* __restrict__ g_mw, double* __restrict__ g_mo,
                               double* __restrict__ g_mg, double* __restrict__ g_rw, double* __restrict__ g_ro,
                               double* __restrict__ g_rg, double* __restrict__ g_pc, double* __restrict__ gtx,
                               double* __restrict__ gty, double* __restrict__ gtz) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  const bool with_pc = F.pc_scale > 0.0;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    const double* T[3] = {tx, ty, tz};
    double* GT[3] = {gtx, gty, gtz};
    const long long st[3] = {G.sx, G.sy, 1};
    const int idx[3] = {i, j, k};
    const int dim[3] = {G.nx, G.ny, G.nz};
    double gmt = 0.0, gmw = 0.0, gmo = 0.0, gmg = 0.0, grw = 0.0, gro = 0.0, grg = 0.0, gpc = 0.0;
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
        const double t = T[ax][lo];
        const long long u = (p0[lo] - p0[hi]) > 0 ? lo : hi;
        const double g_tl = -((lam[lo] - lam[hi]) * (p1[lo] - p1[hi]));
        const double ge = (lam[hi] + g_extra_in[hi]) - (lam[lo] + g_extra_in[lo]);  // extra[hi] - extra[lo]
        double gt = g_tl * mt[u];
        if (u == c) gmt += g_tl * t;
        if (ax == 2 && G.gravity) {
          // gf = T g dz (mw rw + mo ro + mg rg)[u]; extra[lo] -= gf, extra[hi] += gf
          const double k2 = F.g * G.dz;
          gt += ge * k2 * (mw[u] * rw[u] + mo[u] * ro[u] + mg[u] * rg[u]);
          if (u == c) {
            const double s = ge * t * k2;
            gmw += s * rw[c];
            grw += s * mw[c];
            gmo += s * ro[c];
            gro += s * mo[c];
            gmg += s * rg[c];
            grg += s * mg[c];
          }
        }
        if (with_pc) {
          // fpc = T mo[u] (pc[hi] - pc[lo]); extra[lo] -= fpc, extra[hi] += fpc
          const double dpc = pc[hi] - pc[lo];
          gt += ge * mo[u] * dpc;
          if (u == c) gmo += ge * t * dpc;
          const double s = ge * t * mo[u];
          gpc += (c == hi) ? s : -s;
        }
        if (side == 0) GT[ax][c] += gt;
      }
    }
    g_mt[c] += gmt;
    g_mw[c] = gmw;
    g_mo[c] = gmo;
    g_mg[c] = gmg;
    g_rw[c] = grw;
    g_ro[c] = gro;
    g_rg[c] = grg;
    g_pc[c] = gpc;
  }
}

// Well terms of the pressure equation (k_wells_pressure) reversed; the bhp_min
// switch is read from the pressure step's report.
__global__ void k_wells_p_vjp(FluidParams F, int nw, const WellDesc* __restrict__ W, const double* __restrict__ rep,
                              const double* __restrict__ p1, const double* __restrict__ lam,
                              const double* __restrict__ mt, double* __restrict__ g_mt, double* __restrict__ g_val) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int w = 0; w < nw; ++w) {
    const WellDesc& d = W[w];
    const int c = d.cell;
    const double* R = rep + kReportHead + kWellSlots * w;
    const double g_diag = -lam[c] * p1[c], g_b = lam[c];
    if (d.control == 0) {
      if (R[kWMode] == 1.0) {  // switched to its bhp_min: diag += wi mt, b += wi mt pmin
        g_mt[c] += d.wi * (g_diag + g_b * (d.bhp_min * 1e6));
      } else {
        g_val[w] += g_b * (F.rate_scale / kDay) * (d.producer ? -1.0 : 1.0);
      }
    } else if (d.control == 1) {
      g_mt[c] += d.wi * (g_diag + g_b * (d.value * 1e6));
      g_val[w] += g_b * (d.wi * mt[c]) * 1e6;
    }
  }
}

// Cell chain of the pressure step at (P0, S0): mobilities, densities, o-w
// capillarity, accumulation acc = PV ct/dt, and b = acc P0 (diag also holds acc).
__global__ void k_props_p_vjp(GridDev G, FluidParams F, PvtDev T, double dt, const double* __restrict__ p0,
                              const double* __restrict__ p1, const double* __restrict__ sw,
                              const double* __restrict__ sg, const double* __restrict__ pv,
                              const double* __restrict__ lam, const double* __restrict__ acc,
                              const double* __restrict__ g_acc_in, const double* __restrict__ g_mt,
                              const double* __restrict__ g_mw, const double* __restrict__ g_mo,
                              const double* __restrict__ g_mg, const double* __restrict__ g_rw,
                              const double* __restrict__ g_ro, const double* __restrict__ g_rg,
                              const double* __restrict__ g_pc, double* __restrict__ g_p0, double* __restrict__ g_sw,
                              double* __restrict__ g_sg) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    typedef Dual<3> D;
    const D P = var<3>(p0[c], 0), W = var<3>(sw[c], 1), S = var<3>(sg[c], 2);
    D kro, krw, krg, muo, muw, mug;
    relperms_d(F, W, S, kro, krw, krg);
    const PvtAt at = T.n > 0 ? pvt_locate(T, P.v) : PvtAt{};
    visc_d(F, T, at, P / 1e6, muo, muw, mug);
    const D mw = krw / muw, mo = kro / muo, mg = krg / mug;
    const double gmt = g_mt[c];
    const double gw = gmt + g_mw[c], go = gmt + g_mo[c], gg = gmt + g_mg[c];
    double out[3];
    for (int q = 0; q < 3; ++q) out[q] = gw * mw.d[q] + go * mo.d[q] + gg * mg.d[q];
    if (G.gravity) {
      D rw, ro, rg;
      dens_d(F, T, P, rw, ro, rg);
      for (int q = 0; q < 3; ++q) out[q] += g_rw[c] * rw.d[q] + g_ro[c] * ro.d[q] + g_rg[c] * rg.d[q];
    }
    if (F.pc_scale > 0.0) {
      const D pcd = pc_ow_d(F, W);
      for (int q = 0; q < 3; ++q) out[q] += g_pc[c] * pcd.d[q];
    }
    // acc: diag (cotangent -lambda P1) and b = acc P0 (cotangent lambda P0), plus the
    // row-balance injection
    const double g_acc = lam[c] * p0[c] - lam[c] * p1[c] + g_acc_in[c];
    const D ct = ct_d(F, T, P, W, S);
    const double k = pv[c] / dt;
    for (int q = 0; q < 3; ++q) out[q] += g_acc * (k * ct.d[q]);
    g_p0[c] += out[0] + lam[c] * acc[c];
    g_sw[c] += out[1];
    g_sg[c] += out[2];
  }
}

// ---------------------------------------------------------------- rock reverse
// T = 2 a b / (a + b + 1e-15) * area / distance with a = k[c], b = k[c + st]:
// dT/da = 2 b (b + 1e-15) / (a + b + 1e-15)^2 * area / distance.
__global__ void k_trans_vjp(GridDev G, const double* __restrict__ kx, const double* __restrict__ ky,
                            const double* __restrict__ kz, const double* __restrict__ gtx,
                            const double* __restrict__ gty, const double* __restrict__ gtz,
                            double* __restrict__ gkx, double* __restrict__ gky, double* __restrict__ gkz) {
  const long long n = (long long)G.nx * G.ny * G.nz;
  CELL_LOOP(c, n) {
    int i, j, k;
    ijk(G, c, i, j, k);
    const double* K[3] = {kx, ky, kz};
    const double* GT[3] = {gtx, gty, gtz};
    double* GK[3] = {gkx, gky, gkz};
    const long long st[3] = {G.sx, G.sy, 1};
    const int idx[3] = {i, j, k};
    const int dim[3] = {G.nx, G.ny, G.nz};
    const double geo[3] = {G.dy * G.dz / G.dx, G.dx * G.dz / G.dy, G.dx * G.dy / G.dz};
    for (int ax = 0; ax < 3; ++ax) {
      double g = 0.0;
      const double a = K[ax][c];
      if (idx[ax] < dim[ax] - 1 && !(ax == 0 && G.env_nx > 0 && (i + 1) % G.env_nx == 0)) {
        const double b = K[ax][c + st[ax]];
        const double s = a + b + 1e-15;
        g += GT[ax][c] * (2 * b * (b + 1e-15) / (s * s)) * geo[ax];
      }
      if (idx[ax] > 0 && !(ax == 0 && G.env_nx > 0 && i % G.env_nx == 0)) {
        const double b = K[ax][c - st[ax]];
        const double s = a + b + 1e-15;
        g += GT[ax][c - st[ax]] * (2 * b * (b + 1e-15) / (s * s)) * geo[ax];
      }
      GK[ax][c] = g;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- Model (adjoint)
void Model::ensure_adjoint() {
  const long long n = N_;
  for (auto* b : {&gtx_, &gty_, &gtz_, &g_div_w_, &g_div_g_, &g_mw_, &g_mo_, &g_mg_, &g_rw_, &g_ro_, &g_rg_,
                  &g_pc_, &g_mt_p_, &g_acc_, &g_extra_})
    b->reserve(n);
  g_val_.reserve(std::max(nw_, 1));
  g_obs_.reserve(3 * std::max(nw_, 1));
  g_ws_.reserve((size_t)kWellIn * std::max(nw_, 1));
}

void Model::adjoint_reset(hipStream_t s) {
  ensure_adjoint();
  const long long n = N_;
  for (auto* b : {&gtx_, &gty_, &gtz_}) launch_fill_f64(b->p, 0.0, n, s);
  launch_fill_f64(g_val_.p, 0.0, std::max(nw_, 1), s);
}

void Model::transport_vjp(double dt, const double* p_prev, const double* p_new, const double* sw, const double* sg,
                          const double* g_sw1, const double* g_sg1, const double* g_obs_host, double* g_p_new,
                          double* g_p_prev, double* g_sw0, double* g_sg0, hipStream_t s) {
  IMPES_REQUIRE(n_env_ == 1, "adjoint: batched environments are not supported");
  IMPES_REQUIRE(F.max_substeps <= 64, "adjoint: at most 64 substeps");
  ensure_adjoint();
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  const int gr = grid_for(N_, kBlock, 1 << 16);
  const long long n = N_;
  // the transport forward at (P1, S0): mobilities, densities, fluxes, well sources,
  // saturation increments and the substep count (report)
  transport_forward(dt, p_prev, p_new, sw, sg, s);
  // pressure-step cotangents injected by the row balance start from zero
  for (auto* b : {&g_mt_p_, &g_acc_, &g_extra_}) launch_fill_f64(b->p, 0.0, n, s);
  // reverse
  k_sat_vjp<<<gr, kBlock, 0, s>>>(G, F, dt, report_.p, sw, sg, pv.p, dsw_.p, dsg_.p, g_sw1, g_sg1, g_sw0, g_sg0,
                                  g_div_w_.p, g_div_g_.p);
  IMPES_LAUNCH_CHECK();
  k_flux_vjp<<<gr, kBlock, 0, s>>>(G, F, p_new, tx.p, ty.p, tz.p, mob_w_.p, mob_g_.p, rho_w_.p, rho_g_.p, pcog_.p,
                                   g_div_w_.p, g_div_g_.p, g_p_new, g_mw_.p, g_mo_.p, g_mg_.p, g_rw_.p, g_rg_.p,
                                   g_pc_.p, gtx_.p, gty_.p, gtz_.p);
  IMPES_LAUNCH_CHECK();
  if (nw_ > 0) {
    IMPES_HIP_CHECK(hipMemcpyAsync(g_obs_.p, g_obs_host, sizeof(double) * 3 * nw_, hipMemcpyHostToDevice, s));
    WellVjpCtx Y{G, F, pvt, dt, nw_, wells_.p, report_.p, p_prev, p_new, sw, sg, pv.p, mob_w_.p, mob_o_.p,
                 mob_g_.p, mob_t.p, ox.p, oy.p, oz.p, acc.p, tx.p, ty.p, tz.p, g_div_w_.p, g_div_g_.p, g_obs_.p};
    k_wells_sat_vjp<<<1, 256, 0, s>>>(Y, g_ws_.p, g_p_new, g_p_prev, g_mw_.p, g_mo_.p, g_mg_.p, g_sw0, g_sg0,
                                      g_val_.p, g_mt_p_.p, g_acc_.p, g_extra_.p, gtx_.p, gty_.p, gtz_.p);
    IMPES_LAUNCH_CHECK();
  }
  k_props_sat_vjp<<<gr, kBlock, 0, s>>>(G, F, pvt, p_new, sw, sg, g_mw_.p, g_mo_.p, g_mg_.p, g_rw_.p, g_rg_.p,
                                        g_pc_.p, g_p_new, g_sw0, g_sg0);
  IMPES_LAUNCH_CHECK();
  // the host copy of g_obs must outlive the upload
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
}

void Model::solve_adjoint(int precision, double rtol, int restart, int maxit, const double* g, double* lam,
                          SolveStats& st, hipStream_t s) {
  IMPES_REQUIRE(amg_ != nullptr, "adjoint solve before the first forward solve");
  amg_->update_stencil(diag.p, ox.p, oy.p, oz.p, s);
  Operator A;
  A.stencil = true;
  A.st = Stencil7<double>{grid.nx, grid.ny, grid.nz, diag.p, ox.p, oy.p, oz.p};
  A.dinv = dinv.p;
  A.n = N_;
  adj_fgmres_.solve(A, amg_.get(), precision, g, lam, false, rtol, restart, maxit, st, s);
}

void Model::assemble_vjp(double dt, const double* p_prev, const double* sw, const double* sg, const double* p_new,
                         const double* lam, double* g_p_prev, double* g_sw0, double* g_sg0, hipStream_t s) {
  IMPES_REQUIRE(n_env_ == 1, "adjoint: batched environments are not supported");
  ensure_adjoint();
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  const int gr = grid_for(N_, kBlock, 1 << 16);
  // the pressure step's properties at (P0, S0) (the transport overwrote the
  // shared mobility and density buffers)
  pressure_props(dt, p_prev, sw, sg, s);
  k_assemble_vjp<<<gr, kBlock, 0, s>>>(G, F, p_prev, p_new, lam, tx.p, ty.p, tz.p, mob_w_.p, mob_o_.p, mob_g_.p,
                                       mob_t.p, rho_w_.p, rho_o_.p, rho_g_.p, pc_.p, g_extra_.p, g_mt_p_.p, g_mw_.p,
                                       g_mo_.p, g_mg_.p, g_rw_.p, g_ro_.p, g_rg_.p, g_pc_.p, gtx_.p, gty_.p, gtz_.p);
  IMPES_LAUNCH_CHECK();
  if (nw_ > 0) {
    k_wells_p_vjp<<<1, 1, 0, s>>>(F, nw_, wells_.p, report_.p, p_new, lam, mob_t.p, g_mt_p_.p, g_val_.p);
    IMPES_LAUNCH_CHECK();
  }
  k_props_p_vjp<<<gr, kBlock, 0, s>>>(G, F, pvt, dt, p_prev, p_new, sw, sg, pv.p, lam, acc.p, g_acc_.p, g_mt_p_.p,
                                      g_mw_.p, g_mo_.p, g_mg_.p, g_rw_.p, g_ro_.p, g_rg_.p, g_pc_.p, g_p_prev, g_sw0,
                                      g_sg0);
  IMPES_LAUNCH_CHECK();
}

void Model::rock_vjp(const double* kx, const double* ky, const double* kz, double* g_kx, double* g_ky, double* g_kz,
                     hipStream_t s) {
  ensure_adjoint();
  GridDev G = grid_dev(grid, gravity_, env_nx_);
  k_trans_vjp<<<grid_for(N_, kBlock, 1 << 16), kBlock, 0, s>>>(G, kx, ky, kz, gtx_.p, gty_.p, gtz_.p, g_kx, g_ky,
                                                                 g_kz);
  IMPES_LAUNCH_CHECK();
}

void Model::adjoint_values(double* host, hipStream_t s) {
  ensure_adjoint();
  if (nw_ == 0) return;
  IMPES_HIP_CHECK(hipMemcpyAsync(host, g_val_.p, sizeof(double) * nw_, hipMemcpyDeviceToHost, s));
  IMPES_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace impes
