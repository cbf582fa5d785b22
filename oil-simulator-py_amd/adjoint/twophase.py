"""Discrete adjoint of the IMPES step for two-phase oil-water cases.

Physics covered (the subset the history-match case uses): Corey relative
permeabilities, constant viscosities and compressibility (no PVT table),
rate-controlled wells (water injectors with the availability choke, liquid
producers at reservoir rate), gravity on z faces, no capillary pressure. The
forward restates the simulator's step on torch device tensors with the same
arithmetic as the HIP kernels:

  pressure   simulation.py:867-902, 1546-1620, 1622-1660 (TPFA operator with the
             total mobility upwinded by the previous pressure, accumulation
             PV*c_t/dt, rate sources, z-face gravity on the rhs)
  transport  simulation.py:968-1544 (water fluxes upwinded by the new pressure,
             gravity potential on z faces, well sources, dS = dt/PV (q - div),
             N = clip(ceil(max rat), 1, max_substeps) clamped substeps)

The reverse sweep is written by hand (no autograd): the transposed pressure
solve uses the same symmetric operator (A^T = A) and any solver callback (the
HIP AMG-preconditioned FGMRES on the GPU), and every elementwise and stencil
term has its explicit vector-Jacobian product. Piecewise choices (upwind
directions, the substep count N, active clamps, the injector choke) are frozen
at their forward values, i.e. the gradient is the derivative of the smooth
branch the forward took.

Parameters: per-cell log-permeability ln k [ln mD] (isotropic, kx = ky = kz) and
per-well rates [m^3/day]. Observations after each step: producer water cut
fw = lw / (lw + lo + 1e-10) and injector bottom-hole pressure
p_cell + q / (WI * lambda_t).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

MD_TO_M2 = 9.869233e-16  # reservoir.py:35
G = 9.81  # simulation.py:79
DAY = 86400.0
WI = 2.0 * math.pi * 100.0 / math.log(0.28 * math.sqrt(2.0) / 2.0 / 0.1)  # well.py:44-51 at r_w = 0.1


@dataclass
class TwoPhaseCase:
    dims: tuple
    dx: float
    dy: float
    dz: float
    porosity: torch.Tensor          # (nx, ny, nz)
    wells: list                     # dicts: cell (i,j,k), type "injector"|"producer"
    dt: float
    mu_w: float = 1e-3
    mu_o: float = 5e-3
    sw_cr: float = 0.2
    so_r: float = 0.2
    nw: float = 2.0
    no: float = 2.0
    cf: float = 1e-9
    rho_w: float = 1000.0
    rho_o: float = 850.0
    c_w: float = 7.5e-10
    c_o: float = 7.5e-10
    max_ds: float = 0.05
    max_substeps: int = 20
    p0: float = 20e6
    sw0: float = 0.2
    gravity: bool = field(default=True)

    @property
    def pv(self):
        return self.porosity * (self.dx * self.dy * self.dz)

    def flat(self, w) -> int:
        _, ny, nz = self.dims
        i, j, k = w["cell"]
        return (i * ny + j) * nz + k

    @classmethod
    def from_config(cls, config: dict, fields: dict, device="cpu"):
        """The reference-schema config of a two-phase case (no PVT, Corey, rate wells)."""
        rc, fc, sc = config["reservoir"], config["fluid"], config["simulation"]
        dims = tuple(int(v) for v in rc["dimensions"])
        dx, dy, dz = (float(v) for v in rc["grid_size"])
        rp = fc.get("relative_permeability", {})
        if rp.get("model", "corey") != "corey" or fc.get("pvt_table") or fc.get("pvt_path"):
            raise ValueError("the two-phase adjoint covers Corey relative permeabilities without PVT tables")
        if float((fc.get("capillary_pressure", {}) or {}).get("pc_scale", fc.get("pc_scale", 0.0)) or 0.0) != 0.0:
            raise ValueError("the two-phase adjoint has no capillary pressure")
        wells = []
        for w in config["wells"]:
            if w.get("control_type") != "rate" or w.get("rate_type", "reservoir") != "reservoir":
                raise ValueError("the two-phase adjoint needs reservoir-rate wells")
            if w.get("injected_phase", "water") != "water" or w.get("limits") or w.get("bhp_min") is not None:
                raise ValueError("the two-phase adjoint covers water injectors and unlimited producers")
            wells.append({"name": w["name"], "type": w["type"], "cell": (int(w["i"]), int(w["j"]), int(w["k"])),
                          "rate": float(w["control_value"])})
        poro = torch.as_tensor(fields["porosity"], dtype=torch.float64, device=device).reshape(dims)
        return cls(dims=dims, dx=dx, dy=dy, dz=dz, porosity=poro, wells=wells,
                   dt=float(sc["time_step_days"]) * DAY, mu_w=float(fc.get("mu_water", 0.5)) * 1e-3,
                   mu_o=float(fc.get("mu_oil", 1.0)) * 1e-3, sw_cr=float(rp.get("sw_cr", 0.2)),
                   so_r=float(rp.get("so_r", 0.2)), nw=float(rp.get("nw", 2)), no=float(rp.get("no", 2)),
                   cf=(float(fc.get("c_oil", 1e-5)) + float(fc.get("c_water", 1e-5)) + float(fc.get("c_gas", 1e-3))
                       + float(fc.get("c_rock", 1e-5))) / 3 / 1e6,
                   rho_w=float(fc.get("rho_water", 1000.0)), rho_o=float(fc.get("rho_oil", 850.0)),
                   c_w=float(fc.get("c_water", 1e-5)) / 1e6, c_o=float(fc.get("c_oil", 1e-5)) / 1e6,
                   max_ds=float(sc.get("max_saturation_change", 0.05)), max_substeps=int(sc.get("max_substeps", 20)),
                   p0=float(fc.get("pressure", 20.0)) * 1e6, sw0=float(fc.get("s_w", 0.2)),
                   gravity=dims[2] > 1 and dz > 0)


# ---------------------------------------------------------------- pieces
def _harm(a, b, geo):
    return 2 * a * b / (a + b + 1e-15) * geo


def _harm_vjp(a, b, geo, g):
    """d/da and d/db of 2ab/(a+b+eps)*geo times g."""
    den = a + b + 1e-15
    return g * geo * 2 * b * (b + 1e-15) / den ** 2, g * geo * 2 * a * (a + 1e-15) / den ** 2


def transmissibilities(case, lnk):
    """simulation.py:52-74 with kx = ky = kz = exp(ln k) mD."""
    k = torch.exp(lnk) * MD_TO_M2
    dx, dy, dz = case.dx, case.dy, case.dz
    return (_harm(k[:-1], k[1:], dy * dz / dx), _harm(k[:, :-1], k[:, 1:], dx * dz / dy),
            _harm(k[:, :, :-1], k[:, :, 1:], dx * dy / dz))


def transmissibilities_vjp(case, lnk, gx, gy, gz):
    k = torch.exp(lnk) * MD_TO_M2
    dx, dy, dz = case.dx, case.dy, case.dz
    gk = torch.zeros_like(k)
    a, b = _harm_vjp(k[:-1], k[1:], dy * dz / dx, gx)
    gk[:-1] += a; gk[1:] += b
    a, b = _harm_vjp(k[:, :-1], k[:, 1:], dx * dz / dy, gy)
    gk[:, :-1] += a; gk[:, 1:] += b
    a, b = _harm_vjp(k[:, :, :-1], k[:, :, 1:], dx * dy / dz, gz)
    gk[:, :, :-1] += a; gk[:, :, 1:] += b
    return gk * k  # d k / d ln k = k


def mobilities(case, sw):
    """fluid.py:231-248, 377-397 (Corey) over mu -> (lw, lo, lt, dlw/dsw, dlo/dsw)."""
    den = 1 - case.sw_cr - case.so_r
    raw = (sw - case.sw_cr) / den
    sn = torch.clamp(raw, 0.0, 1.0)
    inside = ((raw > 0.0) & (raw < 1.0)).to(sw.dtype)
    krw = sn ** case.nw
    kro = (1 - sn) ** case.no
    dkrw = inside * case.nw * sn ** (case.nw - 1) / den
    dkro = -inside * case.no * (1 - sn) ** (case.no - 1) / den
    lw, lo = krw / case.mu_w, kro / case.mu_o
    return lw, lo, lw + lo, dkrw / case.mu_w, dkro / case.mu_o


def _upw(mask, a):
    """Face values upwinded along the leading axis slice pair (lower, upper)."""
    lo_, hi_ = a
    return torch.where(mask, lo_, hi_)


def _split(t, ax):
    if ax == 0:
        return t[:-1], t[1:]
    if ax == 1:
        return t[:, :-1], t[:, 1:]
    return t[:, :, :-1], t[:, :, 1:]


def _scatter_pair(out, ax, lo_, hi_):
    """out[lower] += lo_; out[upper] += hi_ along axis ax."""
    if ax == 0:
        out[:-1] += lo_; out[1:] += hi_
    elif ax == 1:
        out[:, :-1] += lo_; out[:, 1:] += hi_
    else:
        out[:, :, :-1] += lo_; out[:, :, 1:] += hi_


def _upw_vjp(out, ax, mask, g):
    """VJP of torch.where(mask, a_lower, a_upper) into the cell array `out`."""
    z = torch.zeros_like(g)
    _scatter_pair(out, ax, torch.where(mask, g, z), torch.where(mask, z, g))


def _div(case, f):
    d = torch.zeros(case.dims, dtype=f[0].dtype, device=f[0].device)
    for ax in range(3):
        _scatter_pair(d, ax, f[ax], -f[ax])
    return d


def _div_vjp(g):
    return [a - b for a, b in (_split(g, 0), _split(g, 1), _split(g, 2))]


def _densities(case, p):
    return case.rho_w * (1.0 + case.c_w * (p - 1e5)), case.rho_o * (1.0 + case.c_o * (p - 1e5))


# ---------------------------------------------------------------- forward
class StepTape:
    """What the reverse sweep of one step needs (recomputed from checkpoints)."""


def pressure_system(case, lnk, rates, p, sw, T=None):
    """Face-form operator (diag, ox, oy, oz, all (nx,ny,nz)) and rhs of one step."""
    T = T if T is not None else transmissibilities(case, lnk)
    lw, lo, lt, _, _ = mobilities(case, sw)
    masks = [(a - b) > 0 for a, b in (_split(p, 0), _split(p, 1), _split(p, 2))]
    t = [T[ax] * _upw(masks[ax], _split(lt, ax)) for ax in range(3)]
    acc = case.pv * case.cf / case.dt
    d = torch.zeros_like(p)
    for ax in range(3):
        _scatter_pair(d, ax, t[ax], t[ax])
    diag = d + acc
    off = []
    for ax in range(3):
        o = torch.zeros_like(p)
        if ax == 0: o[:-1] = -t[0]
        elif ax == 1: o[:, :-1] = -t[1]
        else: o[:, :, :-1] = -t[2]
        off.append(o)
    q = torch.zeros_like(p).reshape(-1)
    for w, r in zip(case.wells, rates):
        q[case.flat(w)] += (r / DAY) * (1 if w["type"] == "injector" else -1)
    q = q.reshape(case.dims)
    extra = torch.zeros_like(p)
    if case.gravity:
        rw, ro = _densities(case, p)
        m = masks[2]
        gf = T[2] * G * case.dz * (_upw(m, _split(lw, 2)) * _upw(m, _split(rw, 2))
                                   + _upw(m, _split(lo, 2)) * _upw(m, _split(ro, 2)))
        extra[:, :, :-1] -= gf
        extra[:, :, 1:] += gf
    b = acc * p + q + extra
    return diag, off[0], off[1], off[2], b


def saturation_update(case, lnk, rates, p_new, sw, T=None, tape=None):
    """Transport of one step at the new pressure; returns sw_new (and fills tape)."""
    T = T if T is not None else transmissibilities(case, lnk)
    lw, lo, lt, _, _ = mobilities(case, sw)
    dp = [a - b for a, b in (_split(p_new, 0), _split(p_new, 1), _split(p_new, 2))]
    masks = [x > 0 for x in dp]
    pot = list(dp)
    if case.gravity:
        rw, _ = _densities(case, p_new)
        a, b = _split(rw, 2)
        pot[2] = dp[2] + G * 0.5 * (a + b) * case.dz
    mwf = [_upw(masks[ax], _split(lw, ax)) for ax in range(3)]
    flux = [T[ax] * mwf[ax] * pot[ax] for ax in range(3)]
    div_w = _div(case, flux)
    qw = torch.zeros_like(p_new).reshape(-1)
    pv = case.pv.reshape(-1)
    swf = sw.reshape(-1)
    fw = (lw / (lt + 1e-10)).reshape(-1)
    well_info = []
    for w, r in zip(case.wells, rates):
        idx = case.flat(w)
        q_base = r / DAY
        if w["type"] == "injector":
            up = torch.clamp((1.0 - case.so_r - 0.0) - swf[idx], min=0.0)
            dn = swf[idx] - case.sw_cr
            dS = (case.dt / torch.clamp(pv[idx], min=1e-12)) * q_base
            if dS > 0.0:
                alpha = torch.clamp(up / torch.clamp(dS, min=1e-12), max=1.0)
            else:
                alpha = torch.clamp(dn / torch.clamp(dS.abs(), min=1e-12), max=1.0)
            a = torch.clamp(alpha, 0.0, 1.0)
            qw[idx] += q_base * a
            well_info.append({"choked": bool(alpha < 1.0), "up_pos": bool(up > 0.0), "dS_pos": bool(dS > 0.0)})
        else:
            qw[idx] += (-q_base) * fw[idx]
            well_info.append({})
    qw = qw.reshape(case.dims)
    dSw = (case.dt / case.pv) * (qw - div_w)
    sw_up = torch.clamp(1.0 - case.so_r - 0.0 - sw, min=0.0)
    sw_dn = torch.clamp(sw - case.sw_cr, min=0.0)
    eps = 1e-12
    rat_w = torch.where(dSw > 0, dSw / torch.clamp(sw_up, min=eps), (-dSw) / torch.clamp(sw_dn, min=eps))
    rat = torch.maximum(rat_w, dSw.abs() / case.max_ds)
    rmax = float(rat.detach().max())
    rmax = rmax if math.isfinite(rmax) else 0.0
    N = min(max(int(math.ceil(rmax)), 1), case.max_substeps)
    cur = sw
    active = []
    hi = 1.0 - case.so_r - 0.0
    for _ in range(N):
        u = cur + dSw / N
        active.append((u < case.sw_cr) | (u > hi))
        cur = torch.clamp(torch.clamp(u, min=case.sw_cr), max=hi)
    if tape is not None:
        tape.T, tape.masks_new, tape.pot, tape.mwf, tape.dp_new = T, masks, pot, mwf, dp
        tape.N, tape.active, tape.well_info, tape.fw = N, active, well_info, fw
    return cur, N


def step(case, lnk, rates, p, sw, solve, tape=None):
    """One IMPES step: returns (p_new, sw_new). solve(diag, ox, oy, oz, b) -> p_new."""
    T = transmissibilities(case, lnk)
    diag, ox, oy, oz, b = pressure_system(case, lnk, rates, p, sw, T)
    p_new = solve(diag, ox, oy, oz, b)
    sw_new, _ = saturation_update(case, lnk, rates, p_new, sw, T, tape)
    if tape is not None:
        tape.p, tape.sw, tape.p_new, tape.sw_new = p, sw, p_new, sw_new
        tape.op = (diag, ox, oy, oz)
    return p_new, sw_new


# ---------------------------------------------------------------- observations
def observe(case, rates, p, sw):
    """Producer water cut and injector BHP after a step (one value per well)."""
    lw, lo, lt, _, _ = mobilities(case, sw)
    pf, lwf, ltf = p.reshape(-1), lw.reshape(-1), lt.reshape(-1)
    out = []
    for w, r in zip(case.wells, rates):
        idx = case.flat(w)
        if w["type"] == "producer":
            out.append(lwf[idx] / (ltf[idx] + 1e-10))
        else:
            out.append(pf[idx] + (r / DAY) / (WI * ltf[idx]))
    return torch.stack(out)


def observe_vjp(case, rates, p, sw, g):
    """(d obs/d p)^T g, (d obs/d sw)^T g, (d obs/d rates)^T g."""
    lw, lo, lt, dlw, dlo = mobilities(case, sw)
    gp = torch.zeros_like(p).reshape(-1)
    gs = torch.zeros_like(sw).reshape(-1)
    gr = torch.zeros(len(case.wells), dtype=p.dtype, device=p.device)
    lwf, ltf, dlwf, dlof = lw.reshape(-1), lt.reshape(-1), dlw.reshape(-1), dlo.reshape(-1)
    for n, (w, r) in enumerate(zip(case.wells, rates)):
        idx = case.flat(w)
        if w["type"] == "producer":
            den = ltf[idx] + 1e-10
            dfw_dlw = (den - lwf[idx]) / den ** 2
            dfw_dlo = -lwf[idx] / den ** 2
            gs[idx] += g[n] * (dfw_dlw * dlwf[idx] + dfw_dlo * dlof[idx])
        else:
            gp[idx] += g[n]
            q = r / DAY
            gr[n] += g[n] / (DAY * WI * ltf[idx])
            gs[idx] += g[n] * (-q / (WI * ltf[idx] ** 2)) * (dlwf[idx] + dlof[idx])
    return gp.reshape(p.shape), gs.reshape(sw.shape), gr


# ---------------------------------------------------------------- reverse sweep
def step_vjp(case, lnk, rates, tape, gp_new, gsw_new, solve):
    """Reverse of one step. gp_new/gsw_new: total adjoints of the step's outputs.
    Returns (g_p, g_sw, g_lnk, g_rates) for the step's inputs and parameters."""
    T, p, sw, p_new = tape.T, tape.p, tape.sw, tape.p_new
    lw, lo, lt, dlw, dlo = mobilities(case, sw)
    gT = [torch.zeros_like(t) for t in T]
    g_lw = torch.zeros_like(sw)
    g_lo = torch.zeros_like(sw)
    g_rates = torch.zeros(len(case.wells), dtype=sw.dtype, device=sw.device)
    gp_n = gp_new.clone()
    # --- clamped substeps: sw_{k+1} = clamp(sw_k + dSw/N)
    g = gsw_new
    g_d = torch.zeros_like(sw)
    for act in reversed(tape.active):
        g = torch.where(act, torch.zeros_like(g), g)
        g_d = g_d + g
    g_sw = g  # through sw_old -> first substep
    g_dSw = g_d / tape.N
    # --- dSw = dt/PV (qw - div_w)
    g_qw = g_dSw * (case.dt / case.pv)
    g_div = -g_qw
    g_flux = _div_vjp(g_div)
    for ax in range(3):
        gf = g_flux[ax]
        gT[ax] += gf * tape.mwf[ax] * tape.pot[ax]
        _upw_vjp(g_lw, ax, tape.masks_new[ax], gf * T[ax] * tape.pot[ax])
        g_pot = gf * T[ax] * tape.mwf[ax]
        _scatter_pair(gp_n, ax, g_pot, -g_pot)  # dp = p_lower - p_upper
        if ax == 2 and case.gravity:
            half = g_pot * G * 0.5 * case.dz
            g_rw = torch.zeros_like(sw)
            _scatter_pair(g_rw, 2, half, half)
            gp_n += g_rw * case.rho_w * case.c_w
    # --- well sources
    g_qwf = g_qw.reshape(-1)
    g_swf = g_sw.reshape(-1).clone()
    g_lwf = g_lw.reshape(-1)
    g_lof = g_lo.reshape(-1)
    ltf, lwf = lt.reshape(-1), lw.reshape(-1)
    pv = case.pv.reshape(-1)
    for n, (w, r, info) in enumerate(zip(case.wells, rates, tape.well_info)):
        idx = case.flat(w)
        q_base = r / DAY
        if w["type"] == "injector":
            if not info["choked"]:
                g_rates[n] += g_qwf[idx] / DAY
            elif info["dS_pos"] and info["up_pos"]:
                # q = q_base * up / dS = up * PV / dt, up = 0.8 - sw
                g_swf[idx] += -g_qwf[idx] * pv[idx] / case.dt
        else:
            den = ltf[idx] + 1e-10
            fw = lwf[idx] / den
            g_rates[n] += -g_qwf[idx] * fw / DAY
            g_fw = -g_qwf[idx] * q_base
            g_lwf[idx] += g_fw * (den - lwf[idx]) / den ** 2
            g_lof[idx] += g_fw * (-lwf[idx]) / den ** 2
    g_sw = g_swf.reshape(sw.shape)
    # --- pressure: A p_new = b  ->  A mu = gp_n
    diag, ox, oy, oz = tape.op
    mu = solve(diag, ox, oy, oz, gp_n)
    masks_old = [(a - b) > 0 for a, b in (_split(p, 0), _split(p, 1), _split(p, 2))]
    g_lt = torch.zeros_like(sw)
    g_p = mu * (case.pv * case.cf / case.dt)  # b = acc * p + ...
    for ax in range(3):
        m_lo, m_hi = _split(mu, ax)
        pl, ph = _split(p_new, ax)
        g_t = -(m_lo - m_hi) * (pl - ph)
        lt_up = _upw(masks_old[ax], _split(lt, ax))
        gT[ax] += g_t * lt_up
        _upw_vjp(g_lt, ax, masks_old[ax], g_t * T[ax])
    for n, w in enumerate(case.wells):
        g_rates[n] += mu.reshape(-1)[case.flat(w)] / DAY * (1 if w["type"] == "injector" else -1)
    if case.gravity:
        # extra[:, :, :-1] -= gf ; extra[:, :, 1:] += gf ; b += extra
        m_lo, m_hi = _split(mu, 2)
        g_gf = m_hi - m_lo
        rw, ro = _densities(case, p)
        m = masks_old[2]
        lw_u, lo_u = _upw(m, _split(lw, 2)), _upw(m, _split(lo, 2))
        rw_u, ro_u = _upw(m, _split(rw, 2)), _upw(m, _split(ro, 2))
        s = lw_u * rw_u + lo_u * ro_u
        gT[2] += g_gf * G * case.dz * s
        g_s = g_gf * T[2] * G * case.dz
        _upw_vjp(g_lw, 2, m, g_s * rw_u)
        _upw_vjp(g_lo, 2, m, g_s * ro_u)
        g_rw = torch.zeros_like(sw)
        g_ro = torch.zeros_like(sw)
        _upw_vjp(g_rw, 2, m, g_s * lw_u)
        _upw_vjp(g_ro, 2, m, g_s * lo_u)
        g_p += g_rw * case.rho_w * case.c_w + g_ro * case.rho_o * case.c_o
    # lt = lw + lo
    g_lw += g_lt
    g_lo += g_lt
    g_sw = g_sw + g_lw * dlw + g_lo * dlo
    g_lnk = transmissibilities_vjp(case, lnk, *gT)
    return g_p, g_sw, g_lnk, g_rates


# ---------------------------------------------------------------- misfit + gradient
def misfit(pred, data, sigma):
    return (((pred - data) / sigma) ** 2).sum()


def simulate(case, lnk, rates, solve, steps, checkpoint_every=None):
    """Forward run; returns (observations (steps, wells), checkpoints {step: (p, sw)})."""
    p = torch.full(case.dims, case.p0, dtype=torch.float64, device=lnk.device)
    sw = torch.full(case.dims, case.sw0, dtype=torch.float64, device=lnk.device)
    ck = {0: (p, sw)}
    obs = []
    for n in range(steps):
        p, sw = step(case, lnk, rates, p, sw, solve)
        obs.append(observe(case, rates, p, sw))
        if checkpoint_every and (n + 1) % checkpoint_every == 0 and n + 1 < steps:
            ck[n + 1] = (p, sw)
    return torch.stack(obs), ck


def gradient(case, lnk, rates, data, sigma, solve, steps, checkpoint_every=5, solve_adjoint=None):
    """J = sum ((obs - data)/sigma)^2 and (dJ/d lnk, dJ/d rates) by the adjoint
    with checkpoints every `checkpoint_every` steps and recomputation."""
    solve_adjoint = solve_adjoint or solve
    obs, ck = simulate(case, lnk, rates, solve, steps, checkpoint_every)
    J = misfit(obs, data, sigma)
    g_obs = 2 * (obs - data) / sigma ** 2
    g_lnk = torch.zeros_like(lnk)
    g_rates = torch.zeros_like(rates)
    gp = torch.zeros(case.dims, dtype=torch.float64, device=lnk.device)
    gsw = torch.zeros_like(gp)
    starts = sorted(ck)
    for si in reversed(range(len(starts))):
        s0 = starts[si]
        s1 = starts[si + 1] if si + 1 < len(starts) else steps
        p, sw = ck[s0]
        tapes = []
        for n in range(s0, s1):  # recompute the segment
            tp = StepTape()
            p, sw = step(case, lnk, rates, p, sw, solve, tp)
            tapes.append(tp)
        for n in reversed(range(s0, s1)):
            tp = tapes[n - s0]
            a, b, c = observe_vjp(case, rates, tp.p_new, tp.sw_new, g_obs[n])
            gp = gp + a
            gsw = gsw + b
            g_rates = g_rates + c
            gp, gsw, gl, gr = step_vjp(case, lnk, rates, tp, gp, gsw, solve_adjoint)
            g_lnk = g_lnk + gl
            g_rates = g_rates + gr
        del tapes
    return J, g_lnk, g_rates


class HistoryMatchMisfit(torch.autograd.Function):
    """torch.autograd.Function over the whole run: forward J(lnk, rates), backward
    by the hand-written adjoint (checkpointed, recomputing each segment)."""

    @staticmethod
    def forward(ctx, lnk, rates, case, data, sigma, solve, steps, checkpoint_every):
        with torch.no_grad():
            J, g_lnk, g_rates = gradient(case, lnk.detach(), rates.detach(), data, sigma, solve, steps,
                                         checkpoint_every)
        ctx.save_for_backward(g_lnk, g_rates)
        return J

    @staticmethod
    def backward(ctx, gJ):
        g_lnk, g_rates = ctx.saved_tensors
        return gJ * g_lnk, gJ * g_rates, None, None, None, None, None, None


# ---------------------------------------------------------------- solvers
def dense_solve(diag, ox, oy, oz, b):
    """Differentiable FP64 direct solve (small grids): the autograd reference."""
    A = dense_operator(diag, ox, oy, oz)
    return torch.linalg.solve(A, b.reshape(-1, 1)).reshape(b.shape)


def dense_operator(diag, ox, oy, oz):
    nx, ny, nz = diag.shape
    n = nx * ny * nz
    A = torch.diag_embed(diag.reshape(-1))
    idx = torch.arange(n, device=diag.device).reshape(nx, ny, nz)
    for o, sl_lo, sl_hi in ((ox, (slice(0, -1),), (slice(1, None),)),
                            (oy, (slice(None), slice(0, -1)), (slice(None), slice(1, None))),
                            (oz, (slice(None), slice(None), slice(0, -1)), (slice(None), slice(None), slice(1, None)))):
        r = idx[sl_lo].reshape(-1)
        c = idx[sl_hi].reshape(-1)
        v = o[sl_lo].reshape(-1)
        A = A.index_put((r, c), v, accumulate=True).index_put((c, r), v, accumulate=True)
    return A


def device_solver(rtol=1e-12, precision=32, restart=30, maxit=500):
    """Solve callback on the GPU: the HIP RS-AMG hierarchy (set up on the first
    call, numeric update afterwards: the stencil pattern never changes) inside
    FP64 FGMRES. The operator is symmetric, so the adjoint solves use it too."""
    import impes_hip as H

    state = {}

    def solve(diag, ox, oy, oz, b):
        dims = tuple(diag.shape)
        d, x_, y_, z_ = (t.contiguous().reshape(-1) for t in (diag, ox, oy, oz))
        if state.get("dims") != dims:
            state["h"] = H.AmgHierarchy()
            state["h"].setup_st7(dims, d, x_, y_, z_)
            state["f"] = H.FgmresSolver()
            state["dims"] = dims
        else:
            state["h"].update_st7(d, x_, y_, z_)
        x = torch.zeros_like(d)
        st = state["f"].solve_st7(state["h"], dims, d, x_, y_, z_, 1.0 / d, b.contiguous().reshape(-1), x,
                                  precision=precision, rtol=rtol, restart=restart, maxit=maxit)
        if not st["converged"]:
            raise RuntimeError(f"adjoint pressure solve did not converge: {st}")
        return x.reshape(dims)

    return solve


def scipy_solver():
    """Direct FP64 sparse solve on the host (CPU tests and reference runs)."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    def solve(diag, ox, oy, oz, b):
        nx, ny, nz = diag.shape
        n = nx * ny * nz
        idx = np.arange(n).reshape(nx, ny, nz)
        rows, cols, vals = [idx.reshape(-1)], [idx.reshape(-1)], [diag.detach().cpu().numpy().reshape(-1)]
        for o, lo, hi in ((ox, idx[:-1], idx[1:]), (oy, idx[:, :-1], idx[:, 1:]), (oz, idx[:, :, :-1], idx[:, :, 1:])):
            on = o.detach().cpu().numpy()
            v = (on[:-1] if o is ox else on[:, :-1] if o is oy else on[:, :, :-1]).reshape(-1)
            rows += [lo.reshape(-1), hi.reshape(-1)]
            cols += [hi.reshape(-1), lo.reshape(-1)]
            vals += [v, v]
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
        x = spl.splu(A.tocsc()).solve(b.detach().cpu().numpy().reshape(-1))
        return torch.as_tensor(x, dtype=b.dtype, device=b.device).reshape(b.shape)

    return solve
