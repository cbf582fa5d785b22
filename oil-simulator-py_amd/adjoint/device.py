"""Adjoint of the three-phase black-oil IMPES run on the HIP kernels: the forward
is the device step of the simulator (impes_model_assemble / _solve_pressure /
_saturation), the reverse sweep the hand-written vector-Jacobian-product
kernels of csrc/adjoint.hip (impes_model_transport_vjp, _solve_adjoint,
_assemble_vjp, _rock_vjp), both on one ``impes_model``.

Objective: a producer-rate misfit J = 0.5 sum ((d - data) / sigma)^2 over every
step's surface oil, water and gas rates of every producer (the report's
observation slots). Parameters: per-cell ln kx [ln mD] (ky, kz in the case's
fixed ratios to kx; porosity fixed) and every well's control value.

Checkpointing: the forward keeps the state (p, sw, sg) at the start of every
segment of ``checkpoint_every`` steps; the reverse sweep recomputes a segment's
states from its checkpoint (solves included, from the same starting points:
with the pressure-error check off, etol = 0, the recomputation repeats the
forward bit for bit), then reverses its steps.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from simulator.reservoir import MD_TO_M2


def _flat(t):
    return t.contiguous().reshape(-1)


class DeviceAdjoint:
    """Forward and reverse sweep of the IMPES run of one case on the GPU.

    solver: dict with precision (32/64), rtol (FGMRES relative residual), etol
    (pressure-error check of the forward solves, 0 = off), restart, maxit and
    adj_rtol (relative residual of the transposed solves)."""

    def __init__(self, config: dict, fields: dict | None = None, solver: dict | None = None, device="cuda"):
        from simulator.cases import build

        self.sim = build(config, fields, device)
        self.model = self.sim._model
        self.dims = tuple(int(v) for v in config["reservoir"]["dimensions"])
        self.n = int(np.prod(self.dims))
        self.dt = float(config.get("simulation", {}).get("time_step_days", 1.0)) * 86400.0
        s = dict({"precision": 32, "rtol": 1e-10, "etol": 0.0, "restart": 30, "maxit": 400, "adj_rtol": 1e-10},
                 **(solver or {}))
        self.solver = s
        res = self.sim.reservoir
        kx = _flat(res.permeability_x)
        self.ky_ratio = _flat(res.permeability_y) / kx
        self.kz_ratio = _flat(res.permeability_z) / kx
        self.pv = _flat(self.sim.porous_volume).contiguous()
        self.lnk0 = torch.log(kx / MD_TO_M2)
        self.wells = self.sim.well_manager.get_wells()
        self.values0 = torch.tensor([float(w.control_value) for w in self.wells], dtype=torch.float64)
        self.producers = [i for i, w in enumerate(self.wells) if w.type != "injector"]
        fl = self.sim.fluid
        self.state0 = tuple(_flat(t).clone() for t in (fl.pressure, fl.s_w, fl.s_g))
        self.timing = {}

    # ------------------------------------------------------------ parameters
    def set_params(self, lnk, values):
        kx = torch.exp(_flat(lnk).to(torch.float64)) * MD_TO_M2
        self._k = (kx, kx * self.ky_ratio, kx * self.kz_ratio)
        self.model.set_rock(*self._k, self.pv)
        for w, v in zip(self.wells, [float(x) for x in values]):
            w.control_value = v
        self.sim.update_wells()

    # ------------------------------------------------------------ forward
    def step(self, p, sw, sg):
        """One device IMPES step; returns (p_new, sw_new, sg_new, observations
        (wells, 3) numpy, substep count, solve stats)."""
        from impes_hip.ops import REPORT_HEAD, WELL_SLOTS

        s = self.solver
        m = self.model
        m.assemble(self.dt, p, sw, sg)
        p_new = torch.empty_like(p)
        st = m.solve_pressure(p, p_new, precision=s["precision"], rtol=s["rtol"], restart=s["restart"],
                              maxit=s["maxit"], etol=s["etol"])
        if not st["converged"]:
            raise RuntimeError(f"forward pressure solve did not converge: {st}")
        sw_new, sg_new = torch.empty_like(sw), torch.empty_like(sg)
        rep = m.saturation(self.dt, p, p_new, sw, sg, sw_new, sg_new)
        obs = np.array([rep[REPORT_HEAD + WELL_SLOTS * w + 12: REPORT_HEAD + WELL_SLOTS * w + 15]
                        for w in range(len(self.wells))]).reshape(len(self.wells), 3)
        return p_new, sw_new, sg_new, obs, int(rep[1]), st

    def observe(self, lnk, values, steps):
        """Observations of every step: (steps, wells, 3)."""
        self.set_params(lnk, values)
        p, sw, sg = (t.clone() for t in self.state0)
        out = []
        for _ in range(steps):
            p, sw, sg, obs, _, _ = self.step(p, sw, sg)
            out.append(obs)
        return np.stack(out)

    # ------------------------------------------------------------ gradient
    def gradient(self, lnk, values, steps, data, sigma, checkpoint_every=4):
        """J, dJ/d lnk (device tensor, dims) and dJ/d values (numpy) of the
        producer-rate misfit against data (steps, wells, 3) with weights sigma."""
        data = np.asarray(data, dtype=np.float64)
        sigma = np.asarray(sigma, dtype=np.float64)
        m = self.model
        s = self.solver
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.set_params(lnk, values)
        k = max(1, int(checkpoint_every or steps))
        p, sw, sg = (t.clone() for t in self.state0)
        ckpt, obs, counts = {}, [], []
        for t in range(steps):
            if t % k == 0:
                ckpt[t] = (p.clone(), sw.clone(), sg.clone())
            p, sw, sg, o, N, _ = self.step(p, sw, sg)
            obs.append(o)
            counts.append(N)
        obs = np.stack(obs)
        resid = (obs - data) / sigma
        J = 0.5 * float(np.sum(resid * resid))
        g_obs = resid / sigma
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # reverse sweep
        m.adjoint_reset()
        g_p = torch.zeros(self.n, dtype=torch.float64, device=p.device)
        g_sw = torch.zeros_like(g_p)
        g_sg = torch.zeros_like(g_p)
        lam = torch.empty_like(g_p)
        adj_its = []
        for s0 in sorted(ckpt, reverse=True):
            s1 = min(s0 + k, steps)
            # recompute the segment: the state at the start of each of its steps and
            # the pressure each step solves for
            states, p_new = [ckpt[s0]], []
            for t in range(s0, s1):
                pn, swn, sgn, _, _, _ = self.step(*states[-1])
                p_new.append(pn)
                if t < s1 - 1:
                    states.append((pn, swn, sgn))
            for t in range(s1 - 1, s0 - 1, -1):
                p0, sw0, sg0 = states[t - s0]
                p1 = p_new[t - s0]
                m.assemble(self.dt, p0, sw0, sg0)  # step t's system (and well report)
                g_p_prev = torch.zeros_like(g_p)
                g_sw0 = torch.empty_like(g_p)
                g_sg0 = torch.empty_like(g_p)
                m.transport_vjp(self.dt, p0, p1, sw0, sg0, g_sw, g_sg, g_obs[t], g_p, g_p_prev, g_sw0, g_sg0)
                st = m.solve_adjoint(g_p, lam, precision=s["precision"], rtol=s["adj_rtol"], restart=s["restart"],
                                     maxit=s["maxit"])
                if not st["converged"] and float(g_p.abs().max()) > 0.0:
                    raise RuntimeError(f"adjoint pressure solve did not converge: {st}")
                adj_its.append(int(st["iterations"]))
                m.assemble_vjp(self.dt, p0, sw0, sg0, p1, lam, g_p_prev, g_sw0, g_sg0)
                g_p, g_sw, g_sg = g_p_prev, g_sw0, g_sg0
        gkx, gky, gkz = m.rock_vjp(*self._k)
        kx = self._k[0]
        g_lnk = (gkx + gky * self.ky_ratio + gkz * self.kz_ratio) * kx
        g_val = m.adjoint_values()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        self.timing = {"forward_s": t1 - t0, "reverse_s": t2 - t1, "adjoint_iterations": adj_its[::-1],
                       "substeps": counts}
        return J, g_lnk.reshape(self.dims), g_val, obs


def rate_data(obs_truth, seed, noise=0.02, floor=1e-3):
    """Synthetic data from the truth run's producer rates with `noise` relative
    Gaussian noise, and the weights sigma = noise * max(|data|, floor * the
    largest |rate| of that well and phase over the run)."""
    rng = np.random.default_rng(seed)
    data = obs_truth * (1.0 + noise * rng.standard_normal(obs_truth.shape))
    scale = np.max(np.abs(obs_truth), axis=0, keepdims=True)
    sigma = noise * np.maximum(np.abs(data), floor * scale + 1e-12)
    return data, sigma


def history_match_case(dims=None, steps=None, rate_factor=1.0, p0_mpa=None):
    """The adjoint-at-scale problem: the three-phase gas-injection case (fields,
    PVT table and wells of the 1M-cell benchmark case, shrunk by ``dims``) is the
    truth; the gradient is taken at a prior ln k drawn like the truth's (same
    layer medians, same sigma and correlation length) from another seed.
    ``rate_factor`` scales the rate controls, ``p0_mpa`` overrides the initial
    pressure (small smooth test cases). Returns (config, truth fields, ln k truth
    [ln mD], ln k prior)."""
    from simulator.cases import SEEDS, baseline_case, layered_logk

    cfg, fields = baseline_case(2, dims=dims, steps=steps)
    if rate_factor != 1.0:
        for w in cfg["wells"]:
            if w.get("control_type") == "rate":
                w["control_value"] *= rate_factor
    if p0_mpa is not None:
        cfg["fluid"]["pressure"] = float(p0_mpa)
    shape = tuple(cfg["reservoir"]["dimensions"])
    prior = layered_logk(shape, 50.0, 500.0, 1.5, 5.0, SEEDS["prior"], SEEDS["medians"])
    return cfg, fields, np.log(fields["perm_x"]), prior
