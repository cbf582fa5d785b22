"""Adjoint of the three-phase black-oil IMPES step (the physics of the 1M-cell
gas-injection case: PVT tables, Stone II / Corey relative permeabilities,
capillarity, gravity, surface-rate gas injectors and oil producers, BHP wells,
availability chokes, rate limits, clamped substeps).

The forward restates the simulator's step (simulation.py:867-1544; the HIP
kernels of impes_model.hip compute the same arithmetic) on torch FP64 device
tensors so that torch autograd produces the reverse sweep of every
elementwise, stencil and well term. The pressure solve is one autograd node
(``FaceSolve``): its forward calls a solver callback (the HIP RS-AMG in FP32
inside FP64 FGMRES on the GPU), its backward solves the transposed system with
the same callback (the operator is symmetric, A^T = A) and forms the operator
and right-hand-side cotangents: for A x = b, gb = A^-T g and gA_ij = -gb_i x_j.
Piecewise choices (upwind directions, bhp_min switching, chokes and limits that
bind, the substep count N, active clamps) are taken from the forward values,
i.e. the gradient is that of the smooth branch the forward took. Runs are
checkpointed every k steps (torch.utils.checkpoint); each segment is
recomputed, solves included, during the reverse sweep.

Parameters: per-cell ln k [ln mD] of kx, with ky and kz kept in fixed ratio to
kx (the case's permeability fields set the ratios), and the control value of
every well (rate in m3/day or BHP in MPa). Porosity is fixed. Objective used by
the benchmark: oil in place at surface conditions after the run,
sum PV (So / Bo + Sg Rv / Bg) (PV So without a PVT table).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch
import torch.nn.functional as Fn

MD_TO_M2 = 9.869233e-16  # reservoir.py:35
G = 9.81  # simulation.py:79
DAY = 86400.0


def _wi(radius: float) -> float:
    """well.py:44-51: Peaceman index with k = 100 (raw) and dx = dy = 1."""
    ro = 0.28 * math.sqrt(2.0) / 2
    return 2 * math.pi * 100.0 / (math.log(ro / radius) + 1e-12)


class TorchPVT:
    """pvt.py:48-50 / fluid.py:182-188: piecewise-linear interpolation in MPa,
    clamped at the table ends (np.interp), differentiable in pressure."""

    KEYS = {"Bo": "Bo", "Bw": "Bw", "Bg": "Bg", "mu_o": "mu_o_cP", "mu_w": "mu_w_cP", "mu_g": "mu_g_cP",
            "Rs": "Rs_m3m3", "Rv": "Rv_m3m3"}

    def __init__(self, table: dict, device):
        f64 = dict(dtype=torch.float64, device=device)
        self.P = torch.tensor(table["pressure_MPa"], **f64)
        self.cols = {}
        for k, name in self.KEYS.items():
            v = table.get(name)
            self.cols[k] = torch.zeros_like(self.P) if v is None else torch.tensor(v, **f64)

    def __call__(self, key: str, p):
        x = p / 1e6
        P, fp = self.P, self.cols[key]
        k = (torch.searchsorted(P, x.detach().reshape(-1).contiguous(), right=True) - 1).reshape(x.shape)
        k = k.clamp(0, P.numel() - 2)
        x0, x1, y0, y1 = P[k], P[k + 1], fp[k], fp[k + 1]
        y = y0 + (x - x0) * ((y1 - y0) / (x1 - x0))
        y = torch.where(x <= P[0], fp[0], y)
        return torch.where(x >= P[-1], fp[-1], y)


@dataclass
class ThreePhaseCase:
    dims: tuple
    dx: float
    dy: float
    dz: float
    pv: torch.Tensor            # (nx, ny, nz) pore volume, m^3
    ky_ratio: torch.Tensor      # ky / kx
    kz_ratio: torch.Tensor      # kz / kx
    lnk0: torch.Tensor          # ln kx [ln mD] of the case's fields
    wells: list                 # dicts (see from_config)
    rates0: torch.Tensor        # control values of the wells
    fluid: dict
    pvt: TorchPVT | None
    dt: float
    rate_scale: float = 1.0
    max_ds: float = 0.05
    max_substeps: int = 20

    @classmethod
    def from_config(cls, config: dict, fields: dict | None = None, device="cuda"):
        """The reference-schema config (reservoir / fluid / wells / simulation) with
        the field overrides of simulator.cases (perm_x/y/z in mD, porosity)."""
        import numpy as np

        fields = fields or {}
        f64 = dict(dtype=torch.float64, device=device)
        rc, fc, sc = config["reservoir"], config["fluid"], config.get("simulation", {})
        dims = tuple(int(v) for v in rc.get("dimensions", [10, 10, 1]))
        dx, dy, dz = (float(v) for v in rc.get("grid_size", [10.0, 10.0, 10.0]))
        kh = float(rc.get("permeability", 100.0))
        kv = kh * float(rc.get("k_vertical_fraction", 0.1))

        def field(name, default):
            v = fields.get(name)
            return np.full(dims, default) if v is None else np.asarray(v, dtype=np.float64).reshape(dims)

        kx, ky, kz = field("perm_x", kh), field("perm_y", kh), field("perm_z", kv)
        poro = field("porosity", float(rc.get("porosity", 0.2)))
        rp = fc.get("relative_permeability", {})
        pc = fc["capillary_pressure"] if "capillary_pressure" in fc else fc
        pcg = fc.get("capillary_pressure_og", {})
        fl = {
            "p0": float(fc.get("pressure", 20.0)) * 1e6, "sw0": float(fc.get("s_w", 0.2)),
            "sg0": float(fc.get("s_g", 0.0)),
            "mu_o": float(fc.get("mu_oil", 1.0)) * 1e-3, "mu_w": float(fc.get("mu_water", 0.5)) * 1e-3,
            "mu_g": float(fc.get("mu_gas", 0.02)) * 1e-3,
            "rho_o": float(fc.get("rho_oil", 850.0)), "rho_w": float(fc.get("rho_water", 1000.0)),
            "rho_g": float(fc.get("rho_gas", 1.0)),
            "c_o": float(fc.get("c_oil", 1e-5)) / 1e6, "c_w": float(fc.get("c_water", 1e-5)) / 1e6,
            "c_g": float(fc.get("c_gas", 1e-3)) / 1e6, "c_r": float(fc.get("c_rock", 1e-5)) / 1e6,
            "model": rp.get("model", "corey"), "nw": float(rp.get("nw", 2)), "no": float(rp.get("no", 2)),
            "ng": float(rp.get("ng", 2)), "ko_end_w": float(rp.get("ko_end_w", 1.0)),
            "ko_end_g": float(rp.get("ko_end_g", 1.0)), "sw_cr": float(rp.get("sw_cr", 0.2)),
            "so_r": float(rp.get("so_r", 0.2)), "sg_cr": float(rp.get("sg_cr", 0.0)),
            "pc_scale": float(pc.get("pc_scale", 0.0)), "pc_exponent": float(pc.get("pc_exponent", 1.5)),
            "pcog_scale": float(pcg.get("pc_scale", 0.0)), "pcog_exponent": float(pcg.get("pc_exponent", 1.5)),
        }
        fl["now"] = float(rp.get("now", fl["no"]))
        fl["nog"] = float(rp.get("nog", fl["ng"]))
        fl["cf"] = (fl["c_o"] + fl["c_w"] + fl["c_g"] + fl["c_r"]) / 3
        table = fc.get("pvt_table")
        if table is None and (fc.get("pvt_path") or (fc.get("pvt", {}) or {}).get("path")):
            import json

            with open(fc.get("pvt_path") or fc["pvt"]["path"], encoding="utf-8") as f:
                table = json.load(f)
        pvt = TorchPVT(table, device) if table is not None else None
        nx, ny, nz = dims
        wells, values = [], []
        for w in config.get("wells", []):
            i, j, k = w["coordinates"] if "coordinates" in w else (w["i"], w["j"], w["k"])
            ct, cv = ((w["control"]["type"], w["control"]["value"]) if "control" in w
                      else (w["control_type"], w["control_value"]))
            if not (0 <= i < nx and 0 <= j < ny and 0 <= k < nz):
                raise ValueError(f"well {w['name']} outside the grid")
            wells.append({"cell": (int(i) * ny + int(j)) * nz + int(k), "injector": w["type"] == "injector",
                          "control": ct, "wi": _wi(float(w["radius"])),
                          "gas": w.get("injected_phase", "water") == "gas",
                          "surface": w.get("rate_type", "reservoir") == "surface",
                          "surface_phase": (w.get("surface_phase") or "liquid").lower(),
                          "limits": dict(w.get("limits") or {}), "bhp_min": w.get("bhp_min")})
            values.append(float(cv))
        return cls(dims=dims, dx=dx, dy=dy, dz=dz, pv=torch.tensor(poro * dx * dy * dz, **f64),
                   ky_ratio=torch.tensor(ky / kx, **f64), kz_ratio=torch.tensor(kz / kx, **f64),
                   lnk0=torch.tensor(np.log(kx), **f64), wells=wells, rates0=torch.tensor(values, **f64),
                   fluid=fl, pvt=pvt, dt=float(sc.get("time_step_days", 1.0)) * DAY,
                   rate_scale=float(sc.get("global_rate_scale", 1.0)),
                   max_ds=float(sc.get("max_saturation_change", 0.05)), max_substeps=int(sc.get("max_substeps", 20)))

    # ------------------------------------------------------------ properties (fluid.py)
    def relperms(self, sw, sg):
        """fluid.py:231-248, 377-435 -> (kro, krw, krg)."""
        f = self.fluid
        denom = 1 - f["sw_cr"] - f["so_r"]
        s_norm = torch.clamp((sw - f["sw_cr"]) / denom, 0.0, 1.0)
        krw = s_norm ** f["nw"]
        if f["model"] == "stone2":
            swn = torch.clamp((sw - f["sw_cr"]) / (denom + 1e-12), 0.0, 1.0)
            dg = 1.0 - f["sg_cr"] - f["so_r"]
            sgn = torch.clamp((sg - f["sg_cr"]) / (dg + 1e-12), 0.0, 1.0)
            kro = f["ko_end_w"] * (1.0 - swn) ** f["now"] + f["ko_end_g"] * (1.0 - sgn) ** f["nog"] \
                - min(f["ko_end_w"], f["ko_end_g"])
            kro = torch.clamp(kro, min=0.0)
        else:
            kro = (1 - s_norm) ** f["no"]
        sgn2 = torch.clamp(sg / (1.0 - f["sw_cr"] - f["so_r"] + 1e-10), 0.0, 1.0)
        return kro, krw, sgn2 ** f["ng"]

    def pc_ow(self, sw):
        f = self.fluid
        s_norm = torch.clamp((sw - f["sw_cr"]) / (1 - f["sw_cr"] - f["so_r"]), 0.0, 1.0)
        return f["pc_scale"] * (1.0 - s_norm + 1e-6) ** (-f["pc_exponent"])

    def pc_og(self, sg):
        f = self.fluid
        sgn = torch.clamp(sg / (1 - f["sw_cr"] - f["so_r"] + 1e-12), 0.0, 1.0)
        return f["pcog_scale"] * (1.0 - sgn + 1e-6) ** (-f["pcog_exponent"])

    def viscosities(self, p):
        f = self.fluid
        if self.pvt is None:
            return f["mu_o"], f["mu_w"], f["mu_g"]
        return self.pvt("mu_o", p) * 1e-3, self.pvt("mu_w", p) * 1e-3, self.pvt("mu_g", p) * 1e-3

    def ct(self, p, sw, sg):
        """fluid.py:190-222: central difference +-0.1 MPa of the formation volume factors."""
        f = self.fluid
        if self.pvt is None:
            return torch.full_like(p, f["cf"])
        pm = p / 1e6
        c = {}
        for key in ("Bw", "Bo", "Bg"):
            b0 = self.pvt(key, pm * 1e6)
            dB = (self.pvt(key, (pm + 0.1) * 1e6) - self.pvt(key, (pm - 0.1) * 1e6)) / (2 * 0.1)
            c[key] = -(dB / (b0 + 1e-12)) / 1e6
        so = 1.0 - sw - sg
        return f["c_r"] + sw * c["Bw"] + so * c["Bo"] + sg * c["Bg"]

    def densities(self, p):
        f = self.fluid
        if self.pvt is None:
            return (f["rho_w"] * (1.0 + f["c_w"] * (p - 1e5)), f["rho_o"] * (1.0 + f["c_o"] * (p - 1e5)),
                    f["rho_g"] * (1.0 + f["c_g"] * (p - 1e5)))
        bw, bo, rs, bg, rv = (self.pvt(k, p) for k in ("Bw", "Bo", "Rs", "Bg", "Rv"))
        return (f["rho_w"] / (bw + 1e-12), (f["rho_o"] + rs * f["rho_g"]) / (bo + 1e-12),
                (f["rho_g"] + rv * f["rho_o"]) / (bg + 1e-12))

    def initial_state(self):
        f64 = dict(dtype=torch.float64, device=self.pv.device)
        f = self.fluid
        return (torch.full(self.dims, f["p0"], **f64), torch.full(self.dims, f["sw0"], **f64),
                torch.full(self.dims, f["sg0"], **f64))


# ---------------------------------------------------------------- branch record
def _record(branches, mask):
    """Appends (count, sum of positions) of a boolean decision mask: two runs took
    the same branches iff their records are equal (up to hash collisions, which a
    single flipped decision cannot cause)."""
    if branches is None:
        return
    m = mask.detach().reshape(-1).to(torch.int64)
    pos = torch.arange(m.numel(), device=m.device, dtype=torch.int64)
    branches.append((int(m.sum()), int((m * pos).sum())))


def decisions_changed(a: list, b: list) -> int:
    """Number of decision records that differ between two runs' branch records."""
    return sum(x != y for x, y in zip(a, b)) + abs(len(a) - len(b))


class BranchLog(list):
    """A branch record (the list `_record` fills) that also keeps the decisions
    themselves, in order: upwind masks, the bhp_min switch, choke signs and picks,
    the clamp of the choke factor, rate limits, the substep count N and the four
    saturation clamp masks of every substep. A run given ``BranchLog(replay=log)``
    takes log's decisions at those sites instead of evaluating them, i.e. it
    evaluates the smooth branch that contains log's run: central differences of
    two such runs around the recorded point then measure the derivative the
    adjoint computes. Kinks not replayed: the PVT table segments, the relperm and
    capillary clamps, and the 1e-12 denominator floors."""

    def __init__(self, replay: "BranchLog | None" = None):
        super().__init__()
        self.kept = []
        self.replay = replay
        self._next = 0

    def take(self, value):
        if self.replay is not None:
            value = self.replay.kept[self._next]
            self._next += 1
            return value
        self.kept.append(value.detach().clone() if torch.is_tensor(value) else value)
        return value


def _decide(branches, value, tag=None):
    """A piecewise decision (bool mask tensor, bool or int): kept or replayed by a
    BranchLog, and recorded in the branch record (masks by `_record`, the others
    as (tag, value))."""
    if isinstance(branches, BranchLog):
        value = branches.take(value)
    if branches is not None:
        if torch.is_tensor(value):
            _record(branches, value)
        else:
            branches.append((tag, value))
    return value


# ---------------------------------------------------------------- stencil helpers
def _px(t, lo, hi):
    return Fn.pad(t, (0, 0, 0, 0, lo, hi))


def _py(t, lo, hi):
    return Fn.pad(t, (0, 0, lo, hi, 0, 0))


def _pz(t, lo, hi):
    return Fn.pad(t, (lo, hi, 0, 0, 0, 0))


def _faces(t):
    return t[:-1], t[1:], t[:, :-1], t[:, 1:], t[:, :, :-1], t[:, :, 1:]


def _div(fx, fy, fz):
    """Net outflow per cell of face fluxes (oracle div / k_flux)."""
    return (_px(fx, 0, 1) - _px(fx, 1, 0)) + (_py(fy, 0, 1) - _py(fy, 1, 0)) + (_pz(fz, 0, 1) - _pz(fz, 1, 0))


def transmissibilities(case: ThreePhaseCase, lnk):
    """simulation.py:52-74: harmonic mean with the 1e-15 m^2 epsilon, k from ln k [ln mD]."""
    kx = torch.exp(lnk) * MD_TO_M2
    ky, kz = kx * case.ky_ratio, kx * case.kz_ratio
    Tx = 2 * kx[:-1] * kx[1:] / (kx[:-1] + kx[1:] + 1e-15) * case.dy * case.dz / case.dx
    Ty = 2 * ky[:, :-1] * ky[:, 1:] / (ky[:, :-1] + ky[:, 1:] + 1e-15) * case.dx * case.dz / case.dy
    Tz = 2 * kz[:, :, :-1] * kz[:, :, 1:] / (kz[:, :, :-1] + kz[:, :, 1:] + 1e-15) * case.dx * case.dy / case.dz
    return Tx, Ty, Tz


class FaceSolve(torch.autograd.Function):
    """x = A^-1 b for the symmetric 7-point operator in face form (diag, ox, oy, oz);
    the backward solves with the same operator (A^T = A)."""

    @staticmethod
    def forward(ctx, diag, ox, oy, oz, b, solve):
        x = solve(diag.detach(), ox.detach(), oy.detach(), oz.detach(), b.detach())
        ctx.save_for_backward(diag, ox, oy, oz, x)
        ctx.solve = solve
        return x

    @staticmethod
    def backward(ctx, gx):
        diag, ox, oy, oz, x = ctx.saved_tensors
        lam = ctx.solve(diag.detach(), ox.detach(), oy.detach(), oz.detach(), gx.contiguous())
        gdiag = -lam * x
        gox = _px(-(lam[:-1] * x[1:] + lam[1:] * x[:-1]), 0, 1)
        goy = _py(-(lam[:, :-1] * x[:, 1:] + lam[:, 1:] * x[:, :-1]), 0, 1)
        goz = _pz(-(lam[:, :, :-1] * x[:, :, 1:] + lam[:, :, 1:] * x[:, :, :-1]), 0, 1)
        return gdiag, gox, goy, goz, lam, None


# ---------------------------------------------------------------- pressure step
def pressure_system(case: ThreePhaseCase, T, values, p, sw, sg, branches=None):
    """simulation.py:867-902, 1546-1620, 1622-1660: the face-form operator and rhs
    at the previous state, plus what the transport step needs."""
    f = case.fluid
    Tx, Ty, Tz = T
    kro, krw, krg = case.relperms(sw, sg)
    mu_o, mu_w, mu_g = case.viscosities(p)
    mob_w, mob_o, mob_g = krw / mu_w, kro / mu_o, krg / mu_g
    mob_t = mob_w + mob_o + mob_g
    dpx, dpy, dpz = p[:-1] - p[1:], p[:, :-1] - p[:, 1:], p[:, :, :-1] - p[:, :, 1:]
    ux, uy, uz = (_decide(branches, (d > 0).detach()) for d in (dpx, dpy, dpz))
    xl, xh, yl, yh, zl, zh = _faces(mob_t)
    tx, ty, tz = Tx * torch.where(ux, xl, xh), Ty * torch.where(uy, yl, yh), Tz * torch.where(uz, zl, zh)
    n = p.numel()
    zero = torch.zeros(n, dtype=p.dtype, device=p.device)
    q, q_rate, bhp_diag = zero, zero, zero
    mt, pp = mob_t.reshape(-1), p.reshape(-1)
    terms = []
    for wn, w in enumerate(case.wells):
        c = w["cell"]
        idx = torch.tensor([c], device=p.device)
        if w["control"] == "rate":
            rate = (values[wn] * case.rate_scale) / DAY * (1.0 if w["injector"] else -1.0)
            if not w["injector"] and w["bhp_min"] is not None:
                denom = w["wi"] * mt[c] + 1e-12
                preq = pp[c] - rate / denom
                pmin = float(w["bhp_min"]) * 1e6
                if _decide(branches, float(preq.detach()) < pmin, ("bhp_min", wn)):
                    term = w["wi"] * mt[c]
                    terms.append(term)
                    bhp_diag = bhp_diag.index_add(0, idx, term.reshape(1))
                    q = q.index_add(0, idx, (term * pmin).reshape(1))
                    continue
            terms.append(None)
            q = q.index_add(0, idx, rate.reshape(1))
            q_rate = q_rate.index_add(0, idx, rate.reshape(1))
        elif w["control"] == "bhp":
            term = w["wi"] * mt[c]
            terms.append(term)
            bhp_diag = bhp_diag.index_add(0, idx, term.reshape(1))
            q = q.index_add(0, idx, (term * (values[wn] * 1e6)).reshape(1))
        else:
            terms.append(None)
    acc = case.pv * case.ct(p, sw, sg) / case.dt
    diag = (_px(tx, 0, 1) + _px(tx, 1, 0) + _py(ty, 0, 1) + _py(ty, 1, 0) + _pz(tz, 0, 1) + _pz(tz, 1, 0)
            + acc + bhp_diag.reshape(case.dims))
    ox, oy, oz = _px(-tx, 0, 1), _py(-ty, 0, 1), _pz(-tz, 0, 1)
    extra = torch.zeros_like(p)
    if case.dz > 0 and case.dims[2] > 1:
        rho_w, rho_o, rho_g = case.densities(p)
        up = lambda a: torch.where(uz, a[:, :, :-1], a[:, :, 1:])  # noqa: E731
        gf = Tz * G * case.dz * (up(mob_w) * up(rho_w) + up(mob_o) * up(rho_o) + up(mob_g) * up(rho_g))
        extra = extra - _pz(gf, 0, 1) + _pz(gf, 1, 0)
    if f["pc_scale"] > 0:
        pc = case.pc_ow(sw)
        mxl, mxh, myl, myh, mzl, mzh = _faces(mob_o)
        fx = Tx * torch.where(ux, mxl, mxh) * (pc[1:] - pc[:-1])
        fy = Ty * torch.where(uy, myl, myh) * (pc[:, 1:] - pc[:, :-1])
        fz = Tz * torch.where(uz, mzl, mzh) * (pc[:, :, 1:] - pc[:, :, :-1])
        extra = extra + (_px(fx, 1, 0) - _px(fx, 0, 1)) + (_py(fy, 1, 0) - _py(fy, 0, 1)) \
            + (_pz(fz, 1, 0) - _pz(fz, 0, 1))
    b = acc * p + q.reshape(case.dims) + extra
    return {"diag": diag, "ox": ox, "oy": oy, "oz": oz, "b": b, "acc": acc, "q_rate": q_rate, "extra": extra,
            "terms": terms, "mob_t": mob_t, "tx": tx, "ty": ty, "tz": tz}


# ---------------------------------------------------------------- transport
def _mx(a):
    return torch.clamp(a, min=1e-12)


def _choke(alpha, up, dn, dS, branches=None):
    """min(alpha, up / max(dS, 1e-12)) when dS > 0, else min(alpha, dn / max(|dS|, 1e-12))."""
    pos = _decide(branches, float(dS.detach()) > 0.0, "choke_sign")
    cand = up / _mx(dS) if pos else dn / _mx(torch.abs(dS))
    if _decide(branches, float(cand.detach()) < float(alpha.detach()), "choke_pick"):
        return cand
    return alpha


def _clamp01(alpha, branches=None):
    """clamp(alpha, 0, 1) of a choke factor (alpha <= 1 by construction)."""
    if _decide(branches, float(alpha.detach()) < 0.0, "choke_floor"):
        return torch.zeros_like(alpha)
    return alpha


def saturation_step(case: ThreePhaseCase, T, values, p_prev, p_new, sw_old, sg_old, ps, branches=None, obs=None):
    """simulation.py:968-1544 (row-balance form of the BHP rate): new saturations.
    ``obs`` (a list) receives the step's observations: a (wells, 3) tensor of the
    surface oil, water and gas rates [m3/day] of every producer (rate and BHP
    controlled; formation volume factors 1 without a PVT table), zero rows for
    injectors."""
    f = case.fluid
    Tx, Ty, Tz = T
    P = p_new
    kro, krw, krg = case.relperms(sw_old, sg_old)
    mu_o, mu_w, mu_g = case.viscosities(P)
    mob_w, mob_o, mob_g = krw / mu_w, kro / mu_o, krg / mu_g
    mob_t = mob_w + mob_o + mob_g
    dpx, dpy, dpz = P[:-1] - P[1:], P[:, :-1] - P[:, 1:], P[:, :, :-1] - P[:, :, 1:]
    ux, uy, uz = (_decide(branches, (d > 0).detach()) for d in (dpx, dpy, dpz))
    wl = _faces(mob_w)
    gl = _faces(mob_g)
    mwx, mwy, mwz = torch.where(ux, wl[0], wl[1]), torch.where(uy, wl[2], wl[3]), torch.where(uz, wl[4], wl[5])
    mgx, mgy, mgz = torch.where(ux, gl[0], gl[1]), torch.where(uy, gl[2], gl[3]), torch.where(uz, gl[4], gl[5])
    if case.dz > 0 and case.dims[2] > 1:
        rho_w, _, rho_g = case.densities(P)
        pot_w = dpz + G * 0.5 * (rho_w[:, :, :-1] + rho_w[:, :, 1:]) * case.dz
        pot_g = dpz + G * 0.5 * (rho_g[:, :, :-1] + rho_g[:, :, 1:]) * case.dz
    else:
        pot_w = pot_g = dpz
    div_w = _div(Tx * mwx * dpx, Ty * mwy * dpy, Tz * mwz * pot_w)
    if f["pcog_scale"] > 0:
        pcg = case.pc_og(sg_old)
        div_g = _div(Tx * mgx * (dpx - (pcg[1:] - pcg[:-1])), Ty * mgy * (dpy - (pcg[:, 1:] - pcg[:, :-1])),
                     Tz * mgz * (pot_g - (pcg[:, :, 1:] - pcg[:, :, :-1])))
    else:
        div_g = _div(Tx * mgx * dpx, Ty * mgy * dpy, Tz * mgz * pot_g)
    fw_c, fg_c, fo_c = mob_w / (mob_t + 1e-10), mob_g / (mob_t + 1e-10), mob_o / (mob_t + 1e-10)
    # row balance of the pressure equation at the new pressure: sum over BHP-type
    # wells of WI lambda (p_bhp - p_cell) (simulation.py:1622-1660 inverted)
    flux = _div(ps["tx"] * dpx, ps["ty"] * dpy, ps["tz"] * dpz)
    balance = (flux + ps["acc"] * (p_new - p_prev) - ps["extra"]).reshape(-1) - ps["q_rate"]
    nterm = {}
    for w, term in zip(case.wells, ps["terms"]):
        if term is not None and float(term.detach()) != 0.0:
            nterm[w["cell"]] = nterm.get(w["cell"], 0) + 1
    n = P.numel()
    zero = torch.zeros(n, dtype=P.dtype, device=P.device)
    q_w, q_g = zero, zero
    flat = lambda t: t.reshape(-1)  # noqa: E731
    Pf, swf, sgf, pvf = flat(P), flat(sw_old), flat(sg_old), flat(case.pv)
    mtf, mt_pres = flat(mob_t), flat(ps["mob_t"])
    fwf, fgf, fof = flat(fw_c), flat(fg_c), flat(fo_c)
    pvt = case.pvt
    so_r, sw_cr, sg_cr = f["so_r"], f["sw_cr"], f["sg_cr"]
    one = torch.ones((), dtype=P.dtype, device=P.device)
    zero1 = torch.zeros((), dtype=P.dtype, device=P.device)
    rows = []

    def surface_rates(q, c):
        """|q| [m3/s reservoir] at cell c -> surface oil, water, gas [m3/day]."""
        qa = torch.abs(q)
        fo_, fw_, fg_ = fof[c], fwf[c], fgf[c]
        if pvt is None:
            return torch.stack([qa * fo_ * DAY, qa * fw_ * DAY, qa * fg_ * DAY])
        bo, bw, bg, rs, rv = (pvt(k, Pf[c]) for k in ("Bo", "Bw", "Bg", "Rs", "Rv"))
        return torch.stack([qa * (fo_ / _mx(bo) + fg_ * rv / _mx(bg)) * DAY, qa * (fw_ / _mx(bw)) * DAY,
                            qa * (fg_ / _mx(bg) + fo_ * rs / _mx(bo)) * DAY])

    for wn, w in enumerate(case.wells):
        rows.append(torch.stack([zero1, zero1, zero1]))
        c = w["cell"]
        idx = torch.tensor([c], device=P.device)
        add = lambda arr, v: arr.index_add(0, idx, v.reshape(1))  # noqa: E731
        if w["control"] == "rate":
            q_base = (values[wn] * case.rate_scale) / DAY
            if w["injector"] and w["surface"] and pvt is not None:
                q_total = q_base * pvt("Bg" if w["gas"] else "Bw", Pf[c])
            else:
                q_total = q_base * one
            q_total = q_total * (1.0 if w["injector"] else -1.0)
            if w["injector"]:
                sw, sg = swf[c], sgf[c]
                if w["gas"]:
                    up, dn = torch.clamp((1.0 - so_r - sw) - sg, min=0.0), sg - sg_cr
                else:
                    up, dn = torch.clamp((1.0 - so_r - sg) - sw, min=0.0), sw - sw_cr
                dS = (case.dt / _mx(pvf[c])) * q_total
                alpha = _choke(one, up, dn, dS, branches)
                q_total = q_total * _clamp01(alpha, branches)
                if w["gas"]:
                    q_g = add(q_g, q_total)
                else:
                    q_w = add(q_w, q_total)
            else:
                fo_, fw_, fg_ = fof[c], fwf[c], fgf[c]
                surface = w["surface"] and pvt is not None
                if surface or pvt is not None:
                    bo, bw, bg, rs, rv = (pvt(k, Pf[c]) for k in ("Bo", "Bw", "Bg", "Rs", "Rv"))
                    if surface:
                        sp_ = w["surface_phase"]
                        if sp_ == "oil":
                            q_total = -(q_base / _mx(fo_ / _mx(bo) + fg_ * rv / _mx(bg)))
                        elif sp_ == "water":
                            q_total = -(q_base * bw) / torch.clamp(fw_, min=1e-8)
                        elif sp_ == "gas":
                            q_total = -(q_base / _mx(fg_ / _mx(bg) + fo_ * rs / _mx(bo)))
                        else:
                            q_total = -(q_base / _mx(fo_ / _mx(bo) + fg_ * rv / _mx(bg) + fw_ / _mx(bw)))
                    sw, sg = swf[c], sgf[c]
                    sw_up, sw_dn = torch.clamp((1.0 - so_r - sg) - sw, min=0.0), torch.clamp(sw - sw_cr, min=0.0)
                    sg_up, sg_dn = torch.clamp((1.0 - so_r - sw) - sg, min=0.0), torch.clamp(sg - sg_cr, min=0.0)
                    dSw = (case.dt / _mx(pvf[c])) * (q_total * fw_)
                    dSg = (case.dt / _mx(pvf[c])) * (q_total * fg_)
                    alpha = _choke(_choke(one, sw_up, sw_dn, dSw, branches), sg_up, sg_dn, dSg, branches)
                    q_total = q_total * _clamp01(alpha, branches)
                    lim = w["limits"]
                    if lim:
                        d_o = fo_ / _mx(bo) + fg_ * rv / _mx(bg)
                        d_w = fw_ / _mx(bw)
                        d_g = fg_ / _mx(bg) + fo_ * rs / _mx(bo)
                        d_l = d_o + d_w
                        cand = []
                        for key, dd in (("wopr", d_o), ("wlpr", d_w), ("wgpr", d_g), ("liqr", d_l)):
                            if key in lim and float(dd.detach()) > 0:
                                cand.append((lim[key] / DAY) / dd)
                        if cand:
                            m = cand[0]
                            for cnd in cand[1:]:
                                m = torch.minimum(m, cnd)
                            limited = _decide(branches, float(torch.abs(q_total).detach()) > float(m.detach()),
                                              ("limit", wn))
                            if limited:
                                q_total = -m
                rows[wn] = surface_rates(q_total, c)
                q_w = add(q_w, q_total * fw_)
                q_g = add(q_g, q_total * fg_)
        elif w["control"] == "bhp":
            if nterm.get(c, 0) == 1:
                q_total = -balance[c] * mtf[c] / mt_pres[c]
            else:
                q_total = w["wi"] * mtf[c] * (Pf[c] - values[wn] * 1e6)
            if w["injector"]:
                if w["gas"]:
                    q_g = add(q_g, -q_total)
                else:
                    q_w = add(q_w, -q_total)
            else:
                rows[wn] = surface_rates(q_total, c)
                q_w = add(q_w, -q_total * fwf[c])
                q_g = add(q_g, -q_total * fgf[c])
    if obs is not None:
        obs.append(torch.stack(rows) if rows else torch.zeros((0, 3), dtype=P.dtype, device=P.device))
    dSw = (case.dt / case.pv) * (q_w.reshape(case.dims) - div_w)
    dSg = (case.dt / case.pv) * (q_g.reshape(case.dims) - div_g)
    with torch.no_grad():  # the substep count (a ceil) is taken from the forward values
        eps = 1e-12
        sw_up = torch.clamp(1.0 - so_r - sg_old - sw_old, min=0.0)
        sw_dn = torch.clamp(sw_old - sw_cr, min=0.0)
        sg_up = torch.clamp(1.0 - so_r - sw_old - sg_old, min=0.0)
        sg_dn = torch.clamp(sg_old - sg_cr, min=0.0)
        rat_w = torch.where(dSw > 0, dSw / torch.clamp(sw_up, min=eps), (-dSw) / torch.clamp(sw_dn, min=eps))
        rat_g = torch.where(dSg > 0, dSg / torch.clamp(sg_up, min=eps), (-dSg) / torch.clamp(sg_dn, min=eps))
        rat = torch.maximum(torch.maximum(rat_w, rat_g),
                            torch.maximum(torch.abs(dSw) / case.max_ds, torch.abs(dSg) / case.max_ds))
        rmax = float(rat.max())
        rmax = rmax if math.isfinite(rmax) else 0.0
        N = min(max(int(math.ceil(rmax)), 1), case.max_substeps)
    N = _decide(branches, N, "substeps")
    sw, sg = sw_old, sg_old
    for _ in range(N):
        # sw = min(max(sw_u, sw_cr), 1 - so_r - sg_u), then sg likewise against the
        # new sw, with the active clamps as (replayable) decisions
        sw_u = sw + dSw / N
        sg_u = sg + dSg / N
        lo = _decide(branches, (sw_u < sw_cr).detach())
        a = torch.where(lo, torch.full_like(sw_u, sw_cr), sw_u)
        cap = 1.0 - so_r - sg_u
        hi = _decide(branches, (a > cap).detach())
        sw = torch.where(hi, cap, a)
        lo = _decide(branches, (sg_u < sg_cr).detach())
        a = torch.where(lo, torch.full_like(sg_u, sg_cr), sg_u)
        cap = 1.0 - so_r - sw
        hi = _decide(branches, (a > cap).detach())
        sg = torch.where(hi, cap, a)
    return sw, sg, N


def step(case: ThreePhaseCase, lnk, values, p, sw, sg, solve, direct: bool = False, branches=None, obs=None):
    """One IMPES step; ``direct``: solve is itself differentiable (dense reference);
    ``branches``: a list that receives the record of every piecewise decision;
    ``obs``: a list that receives the step's producer surface rates."""
    T = transmissibilities(case, lnk)
    ps = pressure_system(case, T, values, p, sw, sg, branches)
    if direct:
        p_new = solve(ps["diag"], ps["ox"], ps["oy"], ps["oz"], ps["b"])
    else:
        p_new = FaceSolve.apply(ps["diag"], ps["ox"], ps["oy"], ps["oz"], ps["b"], solve)
    sw_new, sg_new, N = saturation_step(case, T, values, p, p_new, sw, sg, ps, branches, obs)
    return p_new, sw_new, sg_new, N


def oil_in_place(case: ThreePhaseCase, p, sw, sg):
    """sum PV (So / Bo + Sg Rv / Bg) [surface m^3] (PV So without a PVT table)."""
    so = 1.0 - sw - sg
    if case.pvt is None:
        return (case.pv * so).sum()
    return (case.pv * (so / _mx(case.pvt("Bo", p)) + sg * case.pvt("Rv", p) / _mx(case.pvt("Bg", p)))).sum()


def simulate(case: ThreePhaseCase, lnk, values, solve, steps, checkpoint_every=None, direct=False, branches=None):
    """Forward run from the initial state; returns the final (p, sw, sg) and the
    substep counts. With checkpoint_every, each segment of that many steps keeps
    only its input state and is recomputed in the reverse sweep. ``branches`` (a
    list, runs without checkpointing only) receives the decision record of every
    step: equal records of two runs mean the same smooth branch."""
    from torch.utils.checkpoint import checkpoint

    p, sw, sg = case.initial_state()
    counts = []

    def segment(n, p, sw, sg, lnk, values):
        for _ in range(n):
            p, sw, sg, N = step(case, lnk, values, p, sw, sg, solve, direct, branches)
            counts.append(N)
        return p, sw, sg

    done = 0
    while done < steps:
        k = min(checkpoint_every or steps, steps - done)
        if checkpoint_every and torch.is_grad_enabled():
            p, sw, sg = checkpoint(segment, k, p, sw, sg, lnk, values, use_reentrant=False)
        else:
            p, sw, sg = segment(k, p, sw, sg, lnk, values)
        done += k
    return p, sw, sg, list(counts)  # the reverse sweep's recomputations append to `counts` again


def gradient(case: ThreePhaseCase, lnk, values, solve, steps, checkpoint_every=5, direct=False):
    """J = oil in place after `steps` steps and (dJ/d lnk, dJ/d well values)."""
    lnk = lnk.detach().requires_grad_(True)
    values = values.detach().requires_grad_(True)
    p, sw, sg, counts = simulate(case, lnk, values, solve, steps, checkpoint_every, direct)
    J = oil_in_place(case, p, sw, sg)
    g_lnk, g_values = torch.autograd.grad(J, (lnk, values))
    return J.detach(), g_lnk, g_values, counts


# ---------------------------------------------------------------- history match
def observe(case: ThreePhaseCase, lnk, values, solve, steps, direct=False, branches=None):
    """Producer surface rates of every step: (steps, wells, 3) [m3/day]."""
    p, sw, sg = case.initial_state()
    out = []
    for _ in range(steps):
        p, sw, sg, _ = step(case, lnk, values, p, sw, sg, solve, direct, branches, out)
    return torch.stack(out)


def rate_misfit(d, data, sigma):
    """0.5 sum ((d - data) / sigma)^2 over steps, producers and phases."""
    return 0.5 * (((d - data) / sigma) ** 2).sum()


def misfit_gradient(case: ThreePhaseCase, lnk, values, solve, steps, data, sigma, checkpoint_every=4,
                    direct=False):
    """J = rate_misfit of the producer surface rates against data, and
    (dJ/d lnk, dJ/d well values) by torch autograd; segments of checkpoint_every
    steps keep only their input state and are recomputed (solves included)."""
    from torch.utils.checkpoint import checkpoint

    lnk = lnk.detach().requires_grad_(True)
    values = values.detach().requires_grad_(True)

    def segment(s0, k, p, sw, sg, lnk, values):
        J = torch.zeros((), dtype=p.dtype, device=p.device)
        for t in range(s0, s0 + k):
            out = []
            p, sw, sg, _ = step(case, lnk, values, p, sw, sg, solve, direct, None, out)
            J = J + rate_misfit(out[0], data[t], sigma[t])
        return p, sw, sg, J

    p, sw, sg = case.initial_state()
    J = torch.zeros((), dtype=p.dtype, device=p.device)
    done = 0
    while done < steps:
        k = min(checkpoint_every or steps, steps - done)
        if checkpoint_every:
            p, sw, sg, Js = checkpoint(segment, done, k, p, sw, sg, lnk, values, use_reentrant=False)
        else:
            p, sw, sg, Js = segment(done, k, p, sw, sg, lnk, values)
        J = J + Js
        done += k
    g_lnk, g_values = torch.autograd.grad(J, (lnk, values))
    return J.detach(), g_lnk, g_values
