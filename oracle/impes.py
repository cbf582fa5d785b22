"""ORACLE (test infrastructure only): FP64 numpy restatement of the reference's
IMPES time step, with the pressure system solved exactly (scipy splu).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the checker.

Reference files (file:line cited per function): src/simulator/simulation.py,
fluid.py, pvt.py, reservoir.py, well.py. The reference computes its physics in
float32 (torch default dtype); this restatement keeps every formula and branch
but evaluates in float64. One documented change of form: the rate of a
BHP-controlled well in the transport step, WI*lambda_t*(p_cell - p_bhp)
(simulation.py:1306-1309), is evaluated through the pressure equation's row
balance of that cell, which equals it in exact arithmetic but does not cancel
catastrophically (p_cell - p_bhp is a few ulp of p for the reference's
non-SI well index).

Parity pinning (tests/test_oracle_impes.py): golden vectors written by
tests/golden/make_golden.py from the reference's own functions (float32,
compared at float32 tolerance), closed forms (transmissibility, Peaceman index,
Corey/Stone II, np.interp), invariants (symmetric M-matrix, row sums, zero flux
at uniform pressure, volume ledger) and brute-force loops on tiny grids.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

MD_TO_M2 = 9.869233e-16  # reservoir.py:35
G = 9.81  # simulation.py:79
DAY = 86400.0


# ---------------------------------------------------------------- PVT (pvt.py:18-64)
class PVT:
    KEYS = {"Bo": "Bo", "Bw": "Bw", "Bg": "Bg", "mu_o": "mu_o_cP", "mu_w": "mu_w_cP", "mu_g": "mu_g_cP",
            "Rs": "Rs_m3m3", "Rv": "Rv_m3m3"}

    def __init__(self, data: dict):
        self.P = np.asarray(data["pressure_MPa"], dtype=np.float64)
        self.cols = {}
        for k, name in self.KEYS.items():
            if name == "Rv_m3m3" and data.get(name) is None:
                self.cols[k] = np.zeros_like(self.P)
            else:
                self.cols[k] = np.asarray(data.get(name, []), dtype=np.float64)
            if len(self.cols[k]) != len(self.P):
                raise ValueError(f"PVT: column length mismatch ({name})")

    @classmethod
    def from_path(cls, path):
        with open(path, encoding="utf-8") as f:
            return cls(json.load(f))

    def __call__(self, key: str, p_pa: np.ndarray) -> np.ndarray:
        """fluid.py:182-188 + pvt.py:48-50: np.interp in MPa clamped at the ends."""
        fp = self.cols[key]
        return np.interp(np.asarray(p_pa, dtype=np.float64) / 1e6, self.P, fp, left=fp[0], right=fp[-1])


# ---------------------------------------------------------------- inputs
@dataclass
class Well:
    name: str
    type: str
    i: int
    j: int
    k: int
    radius: float
    control_type: str
    control_value: float
    injected_phase: str = "water"
    rate_type: str = "reservoir"
    surface_phase: str | None = None
    limits: dict = field(default_factory=dict)
    bhp_min: float | None = None

    @property
    def wi(self) -> float:
        """well.py:44-51: Peaceman index with k = 100 (raw) and dx = dy = 1."""
        ro = 0.28 * math.sqrt(1.0 + 1.0) / 2
        return 2 * math.pi * 100.0 / (math.log(ro / self.radius) + 1e-12)


def parse_wells(cfgs) -> list[Well]:
    """well.py:64-112 (both config formats)."""
    out = []
    for w in cfgs:
        if "coordinates" in w:
            i, j, k = w["coordinates"]
        elif all(c in w for c in ("i", "j", "k")):
            i, j, k = w["i"], w["j"], w["k"]
        else:
            raise ValueError(f"bad well coordinates: {w}")
        if "control" in w:
            ct, cv = w["control"]["type"], w["control"]["value"]
        elif "control_type" in w and "control_value" in w:
            ct, cv = w["control_type"], w["control_value"]
        else:
            raise ValueError(f"bad well control: {w}")
        out.append(Well(w["name"], w["type"], int(i), int(j), int(k), float(w["radius"]), ct, float(cv),
                        w.get("injected_phase", "water"), w.get("rate_type", "reservoir"), w.get("surface_phase"),
                        w.get("limits") or {}, w.get("bhp_min")))
    return out


class Case:
    """Reservoir + fluid + wells + simulation parameters of one run
    (reservoir.py:8-58, fluid.py:10-110, simulation.py:19-106)."""

    def __init__(self, config: dict, fields: dict | None = None, pvt: PVT | None = None):
        fields = fields or {}
        rc, fc = config["reservoir"], config["fluid"]
        self.sim = dict(config.get("simulation", {}))
        self.dims = tuple(int(v) for v in rc.get("dimensions", [10, 10, 1]))
        nx, ny, nz = self.dims
        self.dx, self.dy, self.dz = (float(v) for v in rc.get("grid_size", [10.0, 10.0, 10.0]))
        poro = fields.get("porosity")
        self.porosity = (np.full(self.dims, float(rc.get("porosity", 0.2))) if poro is None
                         else np.asarray(poro, dtype=np.float64).reshape(self.dims))
        kh = float(rc.get("permeability", 100.0))
        kv = kh * float(rc.get("k_vertical_fraction", 0.1))
        # fields in mD
        self.kx = (np.full(self.dims, kh) if fields.get("perm_x") is None else np.asarray(fields["perm_x"], float).reshape(self.dims)) * MD_TO_M2
        self.ky = (np.full(self.dims, kh) if fields.get("perm_y") is None else np.asarray(fields["perm_y"], float).reshape(self.dims)) * MD_TO_M2
        self.kz = (np.full(self.dims, kv) if fields.get("perm_z") is None else np.asarray(fields["perm_z"], float).reshape(self.dims)) * MD_TO_M2
        self.cell_volume = self.dx * self.dy * self.dz
        self.pv = self.cell_volume * self.porosity
        # fluid.py:26-110
        self.p0 = float(fc.get("pressure", 20.0)) * 1e6
        self.sw0 = float(fc.get("s_w", 0.2))
        self.sg0 = float(fc.get("s_g", 0.0))
        self.mu_o = float(fc.get("mu_oil", 1.0)) * 1e-3
        self.mu_w = float(fc.get("mu_water", 0.5)) * 1e-3
        self.mu_g = float(fc.get("mu_gas", 0.02)) * 1e-3
        self.rho_o_ref = float(fc.get("rho_oil", 850.0))
        self.rho_w_ref = float(fc.get("rho_water", 1000.0))
        self.rho_g_ref = float(fc.get("rho_gas", 1.0))
        self.c_o = float(fc.get("c_oil", 1e-5)) / 1e6
        self.c_w = float(fc.get("c_water", 1e-5)) / 1e6
        self.c_g = float(fc.get("c_gas", 1e-3)) / 1e6
        self.c_r = float(fc.get("c_rock", 1e-5)) / 1e6
        self.cf = (self.c_o + self.c_w + self.c_g + self.c_r) / 3
        if pvt is not None:
            self.pvt = pvt
        else:
            path = fc.get("pvt_path") or (fc.get("pvt", {}) or {}).get("path")
            self.pvt = PVT.from_path(path) if path else None
        rp = fc.get("relative_permeability", {})
        self.model = rp.get("model", "corey")
        self.nw = float(rp.get("nw", 2))
        self.no = float(rp.get("no", 2))
        self.ng = float(rp.get("ng", 2))
        self.ko_end_w = float(rp.get("ko_end_w", 1.0))
        self.ko_end_g = float(rp.get("ko_end_g", 1.0))
        self.now = float(rp.get("now", self.no))
        self.nog = float(rp.get("nog", self.ng))
        self.sw_cr = float(rp.get("sw_cr", 0.2))
        self.so_r = float(rp.get("so_r", 0.2))
        self.sg_cr = float(rp.get("sg_cr", 0.0))
        pc = fc["capillary_pressure"] if "capillary_pressure" in fc else fc
        self.pc_scale = float(pc.get("pc_scale", 0.0))
        self.pc_exponent = float(pc.get("pc_exponent", 1.5))
        pcg = fc.get("capillary_pressure_og", {})
        self.pcog_scale = float(pcg.get("pc_scale", 0.0))
        self.pcog_exponent = float(pcg.get("pc_exponent", 1.5))
        self.wells = parse_wells(config.get("wells", []))
        for w in self.wells:
            if not (0 <= w.i < nx and 0 <= w.j < ny and 0 <= w.k < nz):
                raise ValueError(f"well {w.name} outside the grid")
        self.transmissibilities()

    # ------------------------------------------------------------ a1
    def transmissibilities(self):
        """simulation.py:52-74 (harmonic mean with the 1e-15 m^2 epsilon)."""
        kx, ky, kz = self.kx, self.ky, self.kz
        self.Tx = 2 * kx[:-1] * kx[1:] / (kx[:-1] + kx[1:] + 1e-15) * self.dy * self.dz / self.dx
        self.Ty = 2 * ky[:, :-1] * ky[:, 1:] / (ky[:, :-1] + ky[:, 1:] + 1e-15) * self.dx * self.dz / self.dy
        self.Tz = 2 * kz[:, :, :-1] * kz[:, :, 1:] / (kz[:, :, :-1] + kz[:, :, 1:] + 1e-15) * self.dx * self.dy / self.dz

    # ------------------------------------------------------------ a2
    def relperms(self, sw, sg):
        """fluid.py:231-248, 377-435 -> (kro, krw, krg)."""
        denom = 1 - self.sw_cr - self.so_r
        s_norm = np.clip((sw - self.sw_cr) / denom, 0.0, 1.0)
        krw = s_norm ** self.nw
        if self.model == "stone2":
            swn = np.clip((sw - self.sw_cr) / (denom + 1e-12), 0.0, 1.0)
            dg = 1.0 - self.sg_cr - self.so_r
            sgn = np.clip((sg - self.sg_cr) / (dg + 1e-12), 0.0, 1.0)
            kro = self.ko_end_w * (1.0 - swn) ** self.now + self.ko_end_g * (1.0 - sgn) ** self.nog \
                - min(self.ko_end_w, self.ko_end_g)
            kro = np.maximum(kro, 0.0)
        else:
            kro = (1 - s_norm) ** self.no
        sgn2 = np.clip(sg / (1.0 - self.sw_cr - self.so_r + 1e-10), 0.0, 1.0)
        krg = sgn2 ** self.ng
        return kro, krw, krg

    def pc_ow(self, sw):
        """fluid.py:273-287."""
        if self.pc_scale == 0.0:
            return np.zeros_like(sw)
        s_norm = np.clip((sw - self.sw_cr) / (1 - self.sw_cr - self.so_r), 0.0, 1.0)
        return self.pc_scale * (1.0 - s_norm + 1e-6) ** (-self.pc_exponent)

    def pc_og(self, sg):
        """fluid.py:288-296."""
        if self.pcog_scale == 0.0:
            return np.zeros_like(sg)
        d = 1 - self.sw_cr - self.so_r
        sgn = np.clip(sg / (d + 1e-12), 0.0, 1.0)
        return self.pcog_scale * (1.0 - sgn + 1e-6) ** (-self.pcog_exponent)

    # ------------------------------------------------------------ a3-a5
    def viscosities(self, p):
        """simulation.py:875-877 / 975-977 (Pa s)."""
        if self.pvt is None:
            return self.mu_o, self.mu_w, self.mu_g
        return self.pvt("mu_o", p) * 1e-3, self.pvt("mu_w", p) * 1e-3, self.pvt("mu_g", p) * 1e-3

    def ct(self, p, sw, sg):
        """fluid.py:190-222 (central difference +-0.1 MPa); constant cf without PVT (fluid.py:47-48)."""
        if self.pvt is None:
            return np.full(self.dims, self.cf)
        pm = p / 1e6
        B = {}
        for key in ("Bw", "Bo", "Bg"):
            B[key] = self.pvt(key, pm * 1e6)
            B[key + "p"] = self.pvt(key, (pm + 0.1) * 1e6)
            B[key + "m"] = self.pvt(key, (pm - 0.1) * 1e6)
        c = {}
        for key in ("Bw", "Bo", "Bg"):
            dB = (B[key + "p"] - B[key + "m"]) / (2 * 0.1)
            c[key] = -(dB / (B[key] + 1e-12)) / 1e6
        so = 1.0 - sw - sg
        return self.c_r + sw * c["Bw"] + so * c["Bo"] + sg * c["Bg"]

    def densities(self, p):
        """fluid.py:129-158, 330-375 -> (rho_w, rho_o, rho_g)."""
        if self.pvt is None:
            return (self.rho_w_ref * (1.0 + self.c_w * (p - 1e5)),
                    self.rho_o_ref * (1.0 + self.c_o * (p - 1e5)),
                    self.rho_g_ref * (1.0 + self.c_g * (p - 1e5)))
        bw, bo, rs, bg, rv = (self.pvt(k, p) for k in ("Bw", "Bo", "Rs", "Bg", "Rv"))
        return (self.rho_w_ref / (bw + 1e-12),
                (self.rho_o_ref + rs * self.rho_g_ref) / (bo + 1e-12),
                (self.rho_g_ref + rv * self.rho_o_ref) / (bg + 1e-12))

    # ------------------------------------------------------------ state
    def initial_state(self) -> dict:
        return {"p": np.full(self.dims, self.p0), "sw": np.full(self.dims, self.sw0),
                "sg": np.full(self.dims, self.sg0)}

    def flat(self, w: Well) -> int:
        """simulation.py:816: z-fastest flat index."""
        _, ny, nz = self.dims
        return (w.i * ny + w.j) * nz + w.k


def upwind(mask, a, b):
    return np.where(mask, a, b)


# ---------------------------------------------------------------- pressure
def well_terms(case: Case, mob_t, p_prev, rate_scale):
    """simulation.py:1622-1660 -> q (N), bhp diagonal (N), per-well mode and bhp-term."""
    N = p_prev.size
    q = np.zeros(N)
    q_rate = np.zeros(N)
    diag = np.zeros(N)
    modes, terms = [], []
    mt, pp = mob_t.reshape(-1), p_prev.reshape(-1)
    for w in case.wells:
        idx = case.flat(w)
        if w.control_type == "rate":
            rate = (w.control_value * rate_scale) / DAY * (1 if w.type == "injector" else -1)
            if w.type == "producer" and w.bhp_min is not None:
                denom = w.wi * mt[idx] + 1e-12
                p_req = pp[idx] - rate / denom
                pmin = w.bhp_min * 1e6
                if p_req < pmin:
                    modes.append("bhp")
                    terms.append(w.wi * mt[idx])
                    diag[idx] += w.wi * mt[idx]
                    q[idx] += w.wi * mt[idx] * pmin
                    continue
            modes.append("rate")
            terms.append(0.0)
            q[idx] += rate
            q_rate[idx] += rate
        elif w.control_type == "bhp":
            modes.append(None)
            terms.append(w.wi * mt[idx])
            diag[idx] += w.wi * mt[idx]
            q[idx] += w.wi * mt[idx] * w.control_value * 1e6
        else:
            modes.append(None)
            terms.append(0.0)
    return q, q_rate, diag, modes, terms


def pressure_system(case: Case, st: dict, dt: float) -> dict:
    """simulation.py:867-902 + _build_pressure_matrix_vectorized (:1546-1580) +
    _build_pressure_rhs (:1582-1620). Returns the matrix, rhs and the pieces the
    transport step and the BHP row balance need."""
    p, sw, sg = st["p"], st["sw"], st["sg"]
    nx, ny, nz = case.dims
    kro, krw, krg = case.relperms(sw, sg)
    mu_o, mu_w, mu_g = case.viscosities(p)
    mob_w, mob_o, mob_g = krw / mu_w, kro / mu_o, krg / mu_g
    mob_t = mob_w + mob_o + mob_g
    dpx, dpy, dpz = p[:-1] - p[1:], p[:, :-1] - p[:, 1:], p[:, :, :-1] - p[:, :, 1:]
    tx = case.Tx * upwind(dpx > 0, mob_t[:-1], mob_t[1:])
    ty = case.Ty * upwind(dpy > 0, mob_t[:, :-1], mob_t[:, 1:])
    tz = case.Tz * upwind(dpz > 0, mob_t[:, :, :-1], mob_t[:, :, 1:])
    rate_scale = float(case.sim.get("global_rate_scale", 1.0))
    q, q_rate, bhp_diag, modes, terms = well_terms(case, mob_t, p, rate_scale)
    ct = case.ct(p, sw, sg)
    acc = (case.pv * ct / dt).reshape(-1)
    # face-form operator: off-diagonals -T*lambda, diagonal sum + acc + wells
    N = p.size
    d = np.zeros(case.dims)
    d[:-1] += tx; d[1:] += tx
    d[:, :-1] += ty; d[:, 1:] += ty
    d[:, :, :-1] += tz; d[:, :, 1:] += tz
    diag = d.reshape(-1) + acc + bhp_diag
    ox = np.zeros(case.dims); ox[:-1] = -tx
    oy = np.zeros(case.dims); oy[:, :-1] = -ty
    oz = np.zeros(case.dims); oz[:, :, :-1] = -tz
    # rhs
    extra = np.zeros(case.dims)
    if case.dz > 0 and nz > 1:
        rho_w, rho_o, rho_g = case.densities(p)
        m = dpz > 0
        gf = case.Tz * G * case.dz * (upwind(m, mob_w[:, :, :-1], mob_w[:, :, 1:]) * upwind(m, rho_w[:, :, :-1], rho_w[:, :, 1:])
                                      + upwind(m, mob_o[:, :, :-1], mob_o[:, :, 1:]) * upwind(m, rho_o[:, :, :-1], rho_o[:, :, 1:])
                                      + upwind(m, mob_g[:, :, :-1], mob_g[:, :, 1:]) * upwind(m, rho_g[:, :, :-1], rho_g[:, :, 1:]))
        extra[:, :, :-1] -= gf
        extra[:, :, 1:] += gf
    if case.pc_scale > 0:
        pc = case.pc_ow(sw)
        fx = case.Tx * upwind(dpx > 0, mob_o[:-1], mob_o[1:]) * (pc[1:] - pc[:-1])
        fy = case.Ty * upwind(dpy > 0, mob_o[:, :-1], mob_o[:, 1:]) * (pc[:, 1:] - pc[:, :-1])
        fz = case.Tz * upwind(dpz > 0, mob_o[:, :, :-1], mob_o[:, :, 1:]) * (pc[:, :, 1:] - pc[:, :, :-1])
        extra[1:] += fx; extra[:-1] -= fx
        extra[:, 1:] += fy; extra[:, :-1] -= fy
        extra[:, :, 1:] += fz; extra[:, :, :-1] -= fz
    b = acc * p.reshape(-1) + q + extra.reshape(-1)
    return {"diag": diag, "ox": ox.reshape(-1), "oy": oy.reshape(-1), "oz": oz.reshape(-1), "b": b,
            "acc": acc, "q": q, "q_rate": q_rate, "extra": extra.reshape(-1), "bhp_diag": bhp_diag,
            "modes": modes, "terms": terms, "mob_t": mob_t, "tx": tx, "ty": ty, "tz": tz, "ct": ct}


def system_csr(case: Case, ps: dict) -> sp.csr_matrix:
    nx, ny, nz = case.dims
    n = nx * ny * nz
    sx, sy = ny * nz, nz
    idx = np.arange(n)
    i, j, k = idx // sx, (idx // nz) % ny, idx % nz
    r, c, v = [idx], [idx], [ps["diag"]]
    for off, arr, ok in ((sx, ps["ox"], i < nx - 1), (sy, ps["oy"], j < ny - 1), (1, ps["oz"], k < nz - 1)):
        a = idx[ok]
        r += [a, a + off]
        c += [a + off, a]
        v += [arr[a], arr[a]]
    A = sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(n, n))
    A.sort_indices()
    return A


def solve_direct(A: sp.csr_matrix, b: np.ndarray) -> np.ndarray:
    """Exact FP64 solve (the oracle for every pressure solver)."""
    return spl.splu(A.tocsc()).solve(b)


def bhp_row_balance(case: Case, ps: dict, p_new: np.ndarray, p_prev: np.ndarray) -> np.ndarray:
    """Sum over BHP-type wells of WI*lambda*(p_bhp - p_cell) for every cell, from
    the pressure-equation row: sum_f T*lambda*(p_c - p_nb) + acc*(p_c - p_prev)
    - q_rate - extra."""
    p = p_new.reshape(case.dims)
    flux = np.zeros(case.dims)
    fx = ps["tx"] * (p[:-1] - p[1:])
    fy = ps["ty"] * (p[:, :-1] - p[:, 1:])
    fz = ps["tz"] * (p[:, :, :-1] - p[:, :, 1:])
    flux[:-1] += fx; flux[1:] -= fx
    flux[:, :-1] += fy; flux[:, 1:] -= fy
    flux[:, :, :-1] += fz; flux[:, :, 1:] -= fz
    return flux.reshape(-1) + ps["acc"] * (p_new.reshape(-1) - p_prev.reshape(-1)) - ps["q_rate"] - ps["extra"]


# ---------------------------------------------------------------- transport
def saturation_step(case: Case, st: dict, p_new: np.ndarray, dt: float, ps: dict, bhp_form: str = "balance") -> dict:
    """simulation.py:968-1544. Returns new saturations, the substep count, the
    per-well diagnostics and the ledger increments."""
    nx, ny, nz = case.dims
    sw_old, sg_old = st["sw"], st["sg"]
    P = p_new.reshape(case.dims)
    kro, krw, krg = case.relperms(sw_old, sg_old)
    mu_o, mu_w, mu_g = case.viscosities(P)
    mob_w, mob_o, mob_g = krw / mu_w, kro / mu_o, krg / mu_g
    mob_t = mob_w + mob_o + mob_g
    dpx, dpy, dpz = P[:-1] - P[1:], P[:, :-1] - P[:, 1:], P[:, :, :-1] - P[:, :, 1:]
    mwx = upwind(dpx > 0, mob_w[:-1], mob_w[1:])
    mwy = upwind(dpy > 0, mob_w[:, :-1], mob_w[:, 1:])
    mwz = upwind(dpz > 0, mob_w[:, :, :-1], mob_w[:, :, 1:])
    mgx = upwind(dpx > 0, mob_g[:-1], mob_g[1:])
    mgy = upwind(dpy > 0, mob_g[:, :-1], mob_g[:, 1:])
    mgz = upwind(dpz > 0, mob_g[:, :, :-1], mob_g[:, :, 1:])
    if case.dz > 0 and nz > 1:
        rho_w, _, rho_g = case.densities(P)
        pot_w = dpz + G * 0.5 * (rho_w[:, :, :-1] + rho_w[:, :, 1:]) * case.dz
        pot_g = dpz + G * 0.5 * (rho_g[:, :, :-1] + rho_g[:, :, 1:]) * case.dz
    else:
        pot_w = dpz
        pot_g = dpz
    fwx, fwy, fwz = case.Tx * mwx * dpx, case.Ty * mwy * dpy, case.Tz * mwz * pot_w
    if case.pcog_scale > 0:
        pcg = case.pc_og(sg_old)
        fgx = case.Tx * mgx * (dpx - (pcg[1:] - pcg[:-1]))
        fgy = case.Ty * mgy * (dpy - (pcg[:, 1:] - pcg[:, :-1]))
        fgz = case.Tz * mgz * (pot_g - (pcg[:, :, 1:] - pcg[:, :, :-1]))
    else:
        fgx, fgy, fgz = case.Tx * mgx * dpx, case.Ty * mgy * dpy, case.Tz * mgz * pot_g

    def div(fx, fy, fz):
        d = np.zeros(case.dims)
        d[:-1] += fx; d[1:] -= fx
        d[:, :-1] += fy; d[:, 1:] -= fy
        d[:, :, :-1] += fz; d[:, :, 1:] -= fz
        return d

    div_w, div_g = div(fwx, fwy, fwz), div(fgx, fgy, fgz)
    q_w = np.zeros(case.dims); q_g = np.zeros(case.dims); q_o = np.zeros(case.dims)
    fw = mob_w / (mob_t + 1e-10); fg = mob_g / (mob_t + 1e-10); fo = mob_o / (mob_t + 1e-10)
    pvt = case.pvt
    rate_scale = float(case.sim.get("global_rate_scale", 1.0))
    comp = {"oil_in": 0.0, "oil_out": 0.0, "gas_in": 0.0, "gas_out": 0.0}
    diag = [dict(last_q_total=None, last_surface_rates=None, last_mode=m) for m in ps["modes"]]
    balance = bhp_row_balance(case, ps, p_new, st["p"])
    nterm = np.zeros(P.size, dtype=int)
    for w, term in zip(case.wells, ps["terms"]):
        if term != 0.0:
            nterm[case.flat(w)] += 1
    mob_t_pres = ps["mob_t"].reshape(-1)
    mx = lambda a: max(a, 1e-12)  # noqa: E731

    def bhp_rate(w, mob_t_now):
        """WI*mob_t_now*(p_cell - p_bhp) with p_cell - p_bhp from the row balance."""
        idx = case.flat(w)
        if bhp_form == "balance" and nterm[idx] == 1:
            return -balance[idx] * mob_t_now / mob_t_pres[idx]
        return w.wi * mob_t_now * (P.reshape(-1)[idx] - w.control_value * 1e6)

    pv_cell = lambda w: float(case.pv[w.i, w.j, w.k])  # noqa: E731
    for wn, w in enumerate(case.wells):
        i, j, k = w.i, w.j, w.k
        if w.control_type == "rate":
            q_base = (w.control_value * rate_scale) / DAY
            if w.type == "injector" and w.rate_type == "surface" and pvt is not None:
                q_total = q_base * float(pvt("Bg" if w.injected_phase == "gas" else "Bw", P[i, j, k]))
            else:
                q_total = q_base
            q_total *= 1 if w.type == "injector" else -1
            if w.type == "injector":
                sw, sg = float(sw_old[i, j, k]), float(sg_old[i, j, k])
                if w.injected_phase == "gas":
                    up = max(0.0, (1.0 - case.so_r - sw) - sg)
                    dn = sg - case.sg_cr
                else:
                    up = max(0.0, (1.0 - case.so_r - sg) - sw)
                    dn = sw - case.sw_cr
                dS = (dt / mx(pv_cell(w))) * q_total
                alpha = 1.0
                alpha = min(alpha, up / mx(dS)) if dS > 0.0 else min(alpha, dn / mx(abs(dS)))
                q_total *= max(0.0, min(1.0, alpha))
                if w.injected_phase == "gas":
                    q_g[i, j, k] += q_total
                    if pvt is not None:
                        bg = float(pvt("Bg", P[i, j, k]))
                        qg_surf = max(q_total, 0.0) / mx(bg) * DAY
                        rv = float(pvt("Rv", P[i, j, k]))
                        comp["gas_in"] += qg_surf * (dt / DAY)
                        comp["oil_in"] += rv * qg_surf * (dt / DAY)
                else:
                    q_w[i, j, k] += q_total
            else:
                fo_, fw_, fg_ = float(fo[i, j, k]), float(fw[i, j, k]), float(fg[i, j, k])
                surface = w.rate_type == "surface" and pvt is not None
                if surface or pvt is not None:
                    bo, bw, bg, rs, rv = (float(pvt(key, P[i, j, k])) for key in ("Bo", "Bw", "Bg", "Rs", "Rv"))
                    if surface:
                        sp_ = (w.surface_phase or "liquid").lower()
                        if sp_ == "oil":
                            q_total = -(q_base / mx(fo_ / mx(bo) + fg_ * rv / mx(bg)))
                        elif sp_ == "water":
                            q_total = -(q_base * bw) / max(fw_, 1e-8)
                        elif sp_ == "gas":
                            q_total = -(q_base / mx(fg_ / mx(bg) + fo_ * rs / mx(bo)))
                        else:
                            q_total = -(q_base / mx(fo_ / mx(bo) + fg_ * rv / mx(bg) + fw_ / mx(bw)))
                    sw, sg = float(sw_old[i, j, k]), float(sg_old[i, j, k])
                    sw_up = max(0.0, (1.0 - case.so_r - sg) - sw)
                    sw_dn = max(0.0, sw - case.sw_cr)
                    sg_up = max(0.0, (1.0 - case.so_r - sw) - sg)
                    sg_dn = max(0.0, sg - case.sg_cr)
                    dSw = (dt / mx(pv_cell(w))) * (q_total * fw_)
                    dSg = (dt / mx(pv_cell(w))) * (q_total * fg_)
                    alpha = 1.0
                    alpha = min(alpha, sw_up / mx(dSw)) if dSw > 0.0 else min(alpha, sw_dn / mx(abs(dSw)))
                    alpha = min(alpha, sg_up / mx(dSg)) if dSg > 0.0 else min(alpha, sg_dn / mx(abs(dSg)))
                    q_total *= max(0.0, min(1.0, alpha))
                    if w.limits:
                        d_o = fo_ / mx(bo) + fg_ * rv / mx(bg)
                        d_w = fw_ / mx(bw)
                        d_g = fg_ / mx(bg) + fo_ * rs / mx(bo)
                        d_l = d_o + d_w
                        cand = []
                        if "wopr" in w.limits and d_o > 0: cand.append((w.limits["wopr"] / DAY) / d_o)
                        if "wlpr" in w.limits and d_w > 0: cand.append((w.limits["wlpr"] / DAY) / d_w)
                        if "wgpr" in w.limits and d_g > 0: cand.append((w.limits["wgpr"] / DAY) / d_g)
                        if "liqr" in w.limits and d_l > 0: cand.append((w.limits["liqr"] / DAY) / d_l)
                        if cand:
                            m = min(cand)
                            if abs(q_total) > m:
                                q_total = -m
                    oil_s = abs(q_total) * (fo_ / mx(bo) + fg_ * rv / mx(bg)) * DAY
                    wat_s = abs(q_total) * (fw_ / mx(bw)) * DAY
                    gas_s = abs(q_total) * (fg_ / mx(bg) + fo_ * rs / mx(bo)) * DAY
                    diag[wn]["last_surface_rates"] = {"oil": oil_s, "water": wat_s, "gas": gas_s, "liquid": oil_s + wat_s}
                    diag[wn]["last_q_total"] = float(q_total)
                    qo_res, qg_res = -q_total * fo_, -q_total * fg_
                    comp["oil_out"] += dt * (max(qo_res, 0.0) / mx(bo) + rv * max(qg_res, 0.0) / mx(bg))
                    comp["gas_out"] += dt * (max(qg_res, 0.0) / mx(bg) + rs * max(qo_res, 0.0) / mx(bo))
                q_w[i, j, k] += q_total * fw_
                q_g[i, j, k] += q_total * fg_
                q_o[i, j, k] += q_total * fo_
        elif w.control_type == "bhp":
            q_total = bhp_rate(w, float(mob_t[i, j, k]))
            if w.type == "injector":
                if w.injected_phase == "gas":
                    q_g[i, j, k] += -q_total
                    if pvt is not None:
                        bg, rv = float(pvt("Bg", P[i, j, k])), float(pvt("Rv", P[i, j, k]))
                        qg_surf = max(-q_total, 0.0) / mx(bg) * DAY
                        comp["gas_in"] += qg_surf * (dt / DAY)
                        comp["oil_in"] += rv * qg_surf * (dt / DAY)
                else:
                    q_w[i, j, k] += -q_total
            else:
                fo_, fw_, fg_ = float(fo[i, j, k]), float(fw[i, j, k]), float(fg[i, j, k])
                q_w[i, j, k] += -q_total * fw_
                q_g[i, j, k] += -q_total * fg_
                q_o[i, j, k] += -q_total * fo_
                if pvt is not None:
                    bo, bg, rs, rv = (float(pvt(key, P[i, j, k])) for key in ("Bo", "Bg", "Rs", "Rv"))
                    qa = abs(q_total)
                    qo_res, qg_res = qa * fo_, qa * fg_
                    comp["oil_out"] += dt * (qo_res / mx(bo) + rv * qg_res / mx(bg))
                    comp["gas_out"] += dt * (qg_res / mx(bg) + rs * qo_res / mx(bo))

    dSw = (dt / case.pv) * (q_w - div_w)
    dSg = (dt / case.pv) * (q_g - div_g)
    max_dS = float(case.sim.get("max_saturation_change", 0.05))
    eps = 1e-12
    sw_up = np.maximum(1.0 - case.so_r - sg_old - sw_old, 0.0)
    sw_dn = np.maximum(sw_old - case.sw_cr, 0.0)
    sg_up = np.maximum(1.0 - case.so_r - sw_old - sg_old, 0.0)
    sg_dn = np.maximum(sg_old - case.sg_cr, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        rat_w = np.where(dSw > 0, dSw / np.maximum(sw_up, eps), (-dSw) / np.maximum(sw_dn, eps))
        rat_g = np.where(dSg > 0, dSg / np.maximum(sg_up, eps), (-dSg) / np.maximum(sg_dn, eps))
    rat = np.maximum(np.maximum(rat_w, rat_g), np.maximum(np.abs(dSw) / max_dS, np.abs(dSg) / max_dS))
    rmax = rat.max()
    rmax = rmax if np.isfinite(rmax) else 0.0
    N = int(math.ceil(rmax))
    N = max(N, 1)
    N = min(N, int(case.sim.get("max_substeps", 20)))
    sw_cur, sg_cur = sw_old.copy(), sg_old.copy()
    clamp_w = 0.0
    clamp_g = 0.0
    for _ in range(N):
        sw_u = sw_cur + dSw / N
        sg_u = sg_cur + dSg / N
        sw_cur = np.minimum(np.maximum(sw_u, case.sw_cr), 1.0 - case.so_r - sg_u)
        sg_cur = np.minimum(np.maximum(sg_u, case.sg_cr), 1.0 - case.so_r - sw_cur)
        clamp_w += math.fsum((case.pv * (sw_cur - sw_u)).ravel())
        clamp_g += math.fsum((case.pv * (sg_cur - sg_u)).ravel())
        if pvt is not None:
            _substep_ledger(case, comp, P, sw_cur, sg_cur, dt / N, bhp_rate, rate_scale)
    # mass ledger (simulation.py:1511-1525)
    rho_w, rho_o, rho_g = case.densities(P)
    dsw_e, dsg_e = sw_cur - sw_old, sg_cur - sg_old
    mw, mg = dsw_e * case.pv * rho_w, dsg_e * case.pv * rho_g
    mo = -(dsw_e + dsg_e) * case.pv * rho_o
    mass = {"water_in": mw.clip(min=0).sum(), "water_out": (-mw.clip(max=0)).sum(),
            "gas_in": mg.clip(min=0).sum(), "gas_out": (-mg.clip(max=0)).sum(),
            "oil_in": mo.clip(min=0).sum(), "oil_out": (-mo.clip(max=0)).sum()}
    vol = {"dV_w": math.fsum((case.pv * dsw_e).ravel()), "dV_g": math.fsum((case.pv * dsg_e).ravel()),
           "q_w": dt * math.fsum(q_w.ravel()), "q_g": dt * math.fsum(q_g.ravel()),
           "clamp_w": clamp_w, "clamp_g": clamp_g}
    return {"sw": sw_cur, "sg": sg_cur, "N": N, "wells": diag, "comp": comp, "mass": mass, "vol": vol,
            "q_w": q_w, "q_g": q_g, "q_o": q_o, "dSw": dSw, "dSg": dSg, "rat_max": rmax}


def _substep_ledger(case, comp, P, sw, sg, dt_sub, bhp_rate, rate_scale):
    """simulation.py:1396-1503: component-balance increments of one substep."""
    pvt = case.pvt
    mu_w, mu_o, mu_g = pvt("mu_w", P), pvt("mu_o", P), pvt("mu_g", P)
    kro, krw, krg = case.relperms(sw, sg)
    mw, mo, mg = krw / (mu_w + 1e-12), kro / (mu_o + 1e-12), krg / (mu_g + 1e-12)
    mt = mw + mo + mg
    fw, fg, fo = mw / (mt + 1e-12), mg / (mt + 1e-12), mo / (mt + 1e-12)
    mx = lambda a: max(a, 1e-12)  # noqa: E731
    for w in case.wells:
        i, j, k = w.i, w.j, w.k
        bo, bw, bg, rs, rv = (float(pvt(key, P[i, j, k])) for key in ("Bo", "Bw", "Bg", "Rs", "Rv"))
        fo_, fw_, fg_ = float(fo[i, j, k]), float(fw[i, j, k]), float(fg[i, j, k])
        if w.control_type == "rate":
            q_base = (w.control_value * rate_scale) / DAY
            if w.type == "injector":
                if w.rate_type == "surface":
                    q = q_base * (bg if w.injected_phase == "gas" else bw)
                else:
                    q = q_base
                if w.injected_phase == "gas":
                    qg_surf = max(q, 0.0) / mx(bg) * DAY
                    comp["gas_in"] += qg_surf * (dt_sub / DAY)
                    comp["oil_in"] += rv * qg_surf * (dt_sub / DAY)
            else:
                if w.rate_type == "surface":
                    sp_ = (w.surface_phase or "liquid").lower()
                    if sp_ == "oil":
                        q = -(q_base / mx(fo_ / mx(bo) + fg_ * rv / mx(bg)))
                    elif sp_ == "water":
                        q = -(q_base * bw) / mx(fw_)
                    elif sp_ == "gas":
                        q = -(q_base / mx(fg_ / mx(bg) + fo_ * rs / mx(bo)))
                    else:
                        q = -(q_base / mx(fo_ / mx(bo) + fg_ * rv / mx(bg) + fw_ / mx(bw)))
                else:
                    q = -q_base
                qo_res, qg_res = max(-q * fo_, 0.0), max(-q * fg_, 0.0)
                comp["oil_out"] += dt_sub * (qo_res / mx(bo) + rv * qg_res / mx(bg))
                comp["gas_out"] += dt_sub * (qg_res / mx(bg) + rs * qo_res / mx(bo))
        elif w.control_type == "bhp":
            q = bhp_rate(w, float(mt[i, j, k]))
            if w.type == "injector":
                if w.injected_phase == "gas":
                    qg_surf = max(-q, 0.0) / mx(bg) * DAY
                    comp["gas_in"] += qg_surf * (dt_sub / DAY)
                    comp["oil_in"] += rv * qg_surf * (dt_sub / DAY)
            else:
                qa = abs(q)
                qo_res, qg_res = qa * fo_, qa * fg_
                comp["oil_out"] += dt_sub * (qo_res / mx(bo) + rv * qg_res / mx(bg))
                comp["gas_out"] += dt_sub * (qg_res / mx(bg) + rs * qo_res / mx(bo))


# ---------------------------------------------------------------- driver
class Simulation:
    """Runs IMPES steps (simulation.py:822-865) with the exact pressure solve."""

    def __init__(self, case: Case, bhp_form: str = "balance"):
        self.case = case
        self.state = case.initial_state()
        self.bhp_form = bhp_form
        self.mass_balance = {ph: {"in": 0.0, "out": 0.0} for ph in ("water", "oil", "gas")}
        self.component_balance = {ph: {"in": 0.0, "out": 0.0, "accum": 0.0} for ph in ("oil", "gas")}
        self.volume = {"dV_w": 0.0, "dV_g": 0.0, "q_w": 0.0, "q_g": 0.0, "clamp_w": 0.0, "clamp_g": 0.0}
        self.history = []

    def step(self, dt: float, solver=None) -> dict:
        case = self.case
        ps = pressure_system(case, self.state, dt)
        A = system_csr(case, ps)
        p_new = (solver or solve_direct)(A, ps["b"]).reshape(case.dims)
        out = saturation_step(case, self.state, p_new, dt, ps, self.bhp_form)
        self.state = {"p": p_new, "sw": out["sw"], "sg": out["sg"]}
        for ph in ("water", "oil", "gas"):
            self.mass_balance[ph]["in"] += float(out["mass"][ph + "_in"])
            self.mass_balance[ph]["out"] += float(out["mass"][ph + "_out"])
        cb = self.component_balance
        cb["oil"]["in"] += out["comp"]["oil_in"]; cb["oil"]["out"] += out["comp"]["oil_out"]
        cb["gas"]["in"] += out["comp"]["gas_in"]; cb["gas"]["out"] += out["comp"]["gas_out"]
        cb["oil"]["accum"] = cb["oil"]["in"] - cb["oil"]["out"]
        cb["gas"]["accum"] = cb["gas"]["in"] - cb["gas"]["out"]
        for k, v in out["vol"].items():
            self.volume[k] += v
        rec = {"N": out["N"], "ps": ps, "A": A, "p": p_new, "sat": out}
        self.history.append({"N": out["N"]})
        return rec

    def volume_balance_error(self) -> dict:
        """Cumulative |sum PV dS - dt sum q - clamp| per phase, over total pore volume."""
        v = self.volume
        pvt = float(self.case.pv.sum())
        return {"water": abs(v["dV_w"] - v["q_w"] - v["clamp_w"]) / pvt,
                "gas": abs(v["dV_g"] - v["q_g"] - v["clamp_g"]) / pvt}
