// Device-side black-oil physics of the IMPES step (FP64), restating
// src/simulator/fluid.py and src/simulator/pvt.py of the reference.
#pragma once

#include "common.h"

namespace impes {

enum PvtKey { kBo = 0, kBw = 1, kBg = 2, kMuO = 3, kMuW = 4, kMuG = 5, kRs = 6, kRv = 7, kPvtKeys = 8 };

// Fluid / rock / scheme parameters (fluid.py:26-110, simulation.py:826-828,1354,1377).
struct FluidParams {
  int stone2;
  double nw, no, ng, ko_end_w, ko_end_g, now_, nog;
  double sw_cr, so_r, sg_cr;
  double mu_o, mu_w, mu_g;            // Pa s (no PVT table)
  double rho_o_ref, rho_w_ref, rho_g_ref;
  double c_o, c_w, c_g, c_r, cf;      // 1/Pa
  double pc_scale, pc_exp, pcog_scale, pcog_exp;
  double g;                           // 9.81
  double max_ds;                      // max_saturation_change
  int max_substeps;
  double rate_scale;                  // global_rate_scale
};

// PVT table (pvt.py): pressure in MPa, 8 columns, slopes precomputed on the host.
struct PvtDev {
  int n = 0;
  const double* P = nullptr;
  const double* col = nullptr;    // 8 x n
  const double* slope = nullptr;  // 8 x (n-1)
};

// One well (well.py:10-55 + the branch flags of simulation.py:1051-1347,1622-1660).
struct WellDesc {
  int cell;                // z-fastest flat index (simulation.py:816)
  int producer;            // 0 injector, 1 producer
  int control;             // 0 rate, 1 bhp, 2 other (ignored)
  int gas;                 // injected phase: 0 water, 1 gas
  int surface;             // rate_type == "surface"
  int surface_phase;       // 0 oil, 1 water, 2 gas, 3 liquid
  int has_bhp_min;
  int lim_mask;            // bit0 wopr, bit1 wlpr, bit2 wgpr, bit3 liqr
  double value;            // m3/day or MPa
  double wi;               // Peaceman index (non-SI, well.py:44-51)
  double bhp_min;          // MPa
  double lim[4];           // m3/day
};

// np.interp(p/1e6, P, col, left=col[0], right=col[-1]); products and sums are
// rounded separately, as numpy does. The table segment depends only on the
// pressure, so it is located once (pvt_locate) and read for every column that
// pressure needs (pvt_at): one binary search per pressure instead of one per
// property.
struct PvtAt {
  double x;  // p / 1e6
  int lo;    // segment (inside the table)
  int side;  // 0 inside, -1 below the first node, 1 at or above the last, 2 NaN
};

__device__ __forceinline__ PvtAt pvt_locate(const PvtDev& T, double p_pa) {
  PvtAt a;
  a.x = p_pa / 1e6;
  a.lo = 0;
  const double* xp = T.P;
  const int n = T.n;
  if (!(a.x >= xp[0])) {
    a.side = (a.x != a.x) ? 2 : -1;
    return a;
  }
  if (a.x >= xp[n - 1]) {
    a.side = 1;
    return a;
  }
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= a.x) lo = mid; else hi = mid;
  }
  a.lo = lo;
  a.side = 0;
  return a;
}

// The same location found by walking from a nearby located pressure (a few table
// nodes away): the segment with xp[lo] <= x < xp[lo + 1] is unique in the
// increasing node list, so the walk ends where the binary search does.
__device__ __forceinline__ PvtAt pvt_locate_near(const PvtDev& T, double p_pa, const PvtAt& near) {
  if (near.side != 0) return pvt_locate(T, p_pa);
  PvtAt a;
  a.x = p_pa / 1e6;
  a.lo = 0;
  const double* xp = T.P;
  const int n = T.n;
  if (!(a.x >= xp[0])) {
    a.side = (a.x != a.x) ? 2 : -1;
    return a;
  }
  if (a.x >= xp[n - 1]) {
    a.side = 1;
    return a;
  }
  int lo = near.lo;
  while (lo > 0 && xp[lo] > a.x) --lo;
  while (lo < n - 2 && xp[lo + 1] <= a.x) ++lo;
  a.lo = lo;
  a.side = 0;
  return a;
}

__device__ __forceinline__ double pvt_at(const PvtDev& T, int key, const PvtAt& a) {
  const double* fp = T.col + key * T.n;
  if (a.side == 2) return a.x;
  if (a.side == -1) return fp[0];
  if (a.side == 1) return fp[T.n - 1];
  return __dadd_rn(__dmul_rn(T.slope[key * (T.n - 1) + a.lo], __dsub_rn(a.x, T.P[a.lo])), fp[a.lo]);
}

__device__ __forceinline__ double pvt_eval(const PvtDev& T, int key, double p_pa) {
  return pvt_at(T, key, pvt_locate(T, p_pa));
}

__device__ __forceinline__ double clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }

// x^e for the curve exponents: the common Corey exponents 2 and 1 as the exact
// product (pow goes through log and exp in FP64); the branch is uniform (e is a
// fluid constant)
__device__ __forceinline__ double powe(double x, double e) {
  if (e == 2.0) return x * x;
  if (e == 1.0) return x;
  return pow(x, e);
}

// fluid.py:231-248, 377-435 -> kro, krw, krg
__device__ __forceinline__ void relperms(const FluidParams& F, double sw, double sg, double& kro, double& krw,
                                         double& krg) {
  const double denom = 1.0 - F.sw_cr - F.so_r;
  const double sn = clamp01((sw - F.sw_cr) / denom);
  krw = powe(sn, F.nw);
  if (F.stone2) {
    const double swn = clamp01((sw - F.sw_cr) / (denom + 1e-12));
    const double dg = 1.0 - F.sg_cr - F.so_r;
    const double sgn = clamp01((sg - F.sg_cr) / (dg + 1e-12));
    const double k = F.ko_end_w * powe(1.0 - swn, F.now_) + F.ko_end_g * powe(1.0 - sgn, F.nog) -
                     fmin(F.ko_end_w, F.ko_end_g);
    kro = fmax(k, 0.0);
  } else {
    kro = powe(1.0 - sn, F.no);
  }
  const double sg2 = clamp01(sg / (1.0 - F.sw_cr - F.so_r + 1e-10));
  krg = powe(sg2, F.ng);
}

// fluid.py:273-287
__device__ __forceinline__ double pc_ow(const FluidParams& F, double sw) {
  if (F.pc_scale == 0.0) return 0.0;
  const double sn = clamp01((sw - F.sw_cr) / (1.0 - F.sw_cr - F.so_r));
  return F.pc_scale * pow(1.0 - sn + 1e-6, -F.pc_exp);
}

// fluid.py:288-296
__device__ __forceinline__ double pc_og(const FluidParams& F, double sg) {
  if (F.pcog_scale == 0.0) return 0.0;
  const double d = 1.0 - F.sw_cr - F.so_r;
  const double sn = clamp01(sg / (d + 1e-12));
  return F.pcog_scale * pow(1.0 - sn + 1e-6, -F.pcog_exp);
}

// phase viscosities in Pa s (simulation.py:875-877)
// `at`: the located pressure p (pvt_locate(T, p)) when the caller has it
__device__ __forceinline__ void viscosities(const FluidParams& F, const PvtDev& T, double p, double& mo, double& mw,
                                            double& mg, const PvtAt* at = nullptr) {
  if (T.n > 0) {
    const PvtAt a = at ? *at : pvt_locate(T, p);
    mo = pvt_at(T, kMuO, a) * 1e-3;
    mw = pvt_at(T, kMuW, a) * 1e-3;
    mg = pvt_at(T, kMuG, a) * 1e-3;
  } else {
    mo = F.mu_o;
    mw = F.mu_w;
    mg = F.mu_g;
  }
}

// fluid.py:190-222 (central difference of B over +-0.1 MPa); fluid.py:47-48 without PVT
// `at`: the located pressure p when the caller has it
__device__ __forceinline__ double total_compressibility(const FluidParams& F, const PvtDev& T, double p, double sw,
                                                        double sg, const PvtAt* at = nullptr) {
  if (T.n == 0) return F.cf;
  const double pm = p / 1e6;
  double c[3];
  const int keys[3] = {kBw, kBo, kBg};
  const double p0 = pm * 1e6;
  const PvtAt a0 = (at && p0 == p) ? *at : pvt_locate(T, p0);
  const PvtAt ap = pvt_locate_near(T, (pm + 0.1) * 1e6, a0), am = pvt_locate_near(T, (pm - 0.1) * 1e6, a0);
  for (int q = 0; q < 3; ++q) {
    const double b0 = pvt_at(T, keys[q], a0);
    const double bp = pvt_at(T, keys[q], ap);
    const double bm = pvt_at(T, keys[q], am);
    const double dB = (bp - bm) / (2 * 0.1);
    c[q] = -(dB / (b0 + 1e-12)) / 1e6;
  }
  const double so = 1.0 - sw - sg;
  return F.c_r + sw * c[0] + so * c[1] + sg * c[2];
}

// fluid.py:129-158, 330-375
__device__ __forceinline__ void densities(const FluidParams& F, const PvtDev& T, double p, double& rw, double& ro,
                                          double& rg, const PvtAt* at = nullptr) {
  if (T.n == 0) {
    rw = F.rho_w_ref * (1.0 + F.c_w * (p - 1e5));
    ro = F.rho_o_ref * (1.0 + F.c_o * (p - 1e5));
    rg = F.rho_g_ref * (1.0 + F.c_g * (p - 1e5));
    return;
  }
  const PvtAt a = at ? *at : pvt_locate(T, p);
  const double bw = pvt_at(T, kBw, a), bo = pvt_at(T, kBo, a), rs = pvt_at(T, kRs, a);
  const double bg = pvt_at(T, kBg, a), rv = pvt_at(T, kRv, a);
  rw = F.rho_w_ref / (bw + 1e-12);
  ro = (F.rho_o_ref + rs * F.rho_g_ref) / (bo + 1e-12);
  rg = (F.rho_g_ref + rv * F.rho_o_ref) / (bg + 1e-12);
}

}  // namespace impes
