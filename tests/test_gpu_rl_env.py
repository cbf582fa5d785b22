"""The reinforcement-learning environments (reference src/rl_env.py:15-154) on the
device IMPES step, and the batched environment against single environments.

Case: the golden "gas3" configuration (20x20x3, two rate injectors, two BHP
producers), pressure solve "amg" to a 1e-10 relative residual so that a copy
inside the batch and a single environment differ only by solver roundoff.
Bars: pressure 1e-8 relative, saturations 1e-6 absolute, rewards 1e-6
relative, substep counts and well controls exact."""
import json

import numpy as np
import pytest
import torch

from test_gpu_reference_api import gas3_config

pytestmark = pytest.mark.gpu

N_WELLS = 4


def _cfg_path(tmp_path):
    cfg = gas3_config()
    cfg["simulation"]["amg"] = {"tol": 1e-10, "max_iter": 400}
    path = tmp_path / "env.json"
    path.write_text(json.dumps(cfg))
    return str(path), cfg


def _oil_mass(env):
    return torch.sum(env.reservoir.porosity * env.fluid.s_o * env.fluid.rho_o * env.reservoir.cell_volume).item()


def test_reservoir_env_api(tmp_path, dev):
    from rl_env import ReservoirEnv

    path, cfg = _cfg_path(tmp_path)
    env = ReservoirEnv(path, max_steps=3, device=dev)
    n = 20 * 20 * 3
    assert env.action_space.shape == (N_WELLS,)
    assert env.observation_space.shape == (2 * n,)
    obs, info = env.reset(seed=0)
    assert info == {} and obs.shape == (2 * n,) and obs.dtype == np.float32
    assert np.allclose(obs[:n], 25e6 / 1e7) and np.allclose(obs[n:], 0.2)
    m0 = _oil_mass(env)
    a = np.array([0.5, -1.0, 0.2, -0.4])
    obs, r, term, trunc, info = env.step(a)
    wells = env.well_manager.wells
    # rl_env.py:132-144: rate wells +-20 % of the current control, BHP wells +-5 MPa
    assert wells[0].control_value == pytest.approx(50.0 * 1.1)
    assert wells[1].control_value == pytest.approx(50.0 * 0.8)
    assert wells[2].control_value == pytest.approx(20.0 + 1.0)
    assert wells[3].control_value == pytest.approx(20.0 - 2.0)
    assert info == {"converged": True} and not term and not trunc
    assert r == pytest.approx((m0 - _oil_mass(env)) / 1e6, rel=1e-12)
    assert r > 0.0  # the producers take oil out
    p = env.fluid.pressure.cpu().numpy().ravel()
    assert np.array_equal(obs[:n], (p / 1e7).astype(np.float32))
    # the action compounds on the current control (the reference keeps no base value)
    env.step(np.array([0.5, 0.0, 0.0, 0.0]))
    assert wells[0].control_value == pytest.approx(50.0 * 1.1 * 1.1)
    obs, r, term, trunc, info = env.step(np.zeros(N_WELLS))
    assert term and not trunc
    obs, info = env.reset()
    assert wells is not env.well_manager.wells and env.well_manager.wells[0].control_value == 50.0


def test_batched_env_matches_single_envs(tmp_path, dev):
    from rl_env import BatchedReservoirEnv, ReservoirEnv

    path, cfg = _cfg_path(tmp_path)
    B, steps = 3, 4
    rng = np.random.default_rng(7)
    actions = rng.uniform(-1.0, 1.0, size=(steps, B, N_WELLS))
    benv = BatchedReservoirEnv(path, B, max_steps=steps, device=dev)
    assert benv.action_space.shape == (B, N_WELLS)
    obs_b, _ = benv.reset(seed=1)
    singles = [ReservoirEnv(path, max_steps=steps, device=dev) for _ in range(B)]
    for s in singles:
        s.reset()
    assert np.array_equal(obs_b, np.stack([s._get_observation() for s in singles]))
    dims = (20, 20, 3)
    for t in range(steps):
        ob, rb, tb, trb, ib = benv.step(actions[t])
        assert ib["converged"] and len(ib["substeps"]) == B
        for e, s in enumerate(singles):
            o, r, term, trunc, info = s.step(actions[t, e])
            assert info["converged"]
            assert ib["substeps"][e] == s.simulator.last_substeps, (t, e)
            assert bool(tb[e]) == term and bool(trb[e]) == trunc
            assert rb[e] == pytest.approx(r, rel=1e-6), (t, e)
            pb = benv.fluid.pressure.view(B, *dims)[e]
            ps = s.fluid.pressure
            assert (pb - ps).abs().max().item() <= 1e-8 * ps.abs().max().item(), (t, e)
            for name in ("s_w", "s_g", "s_o"):
                fb = getattr(benv.fluid, name).view(B, *dims)[e]
                assert (fb - getattr(s.fluid, name)).abs().max().item() <= 1e-6, (t, e, name)
            ctrl_b = [w.control_value for w in benv.well_manager.wells[e * N_WELLS:(e + 1) * N_WELLS]]
            assert ctrl_b == [w.control_value for w in s.well_manager.wells]
            assert np.allclose(ob[e], o, rtol=0, atol=2e-6)
    assert all(tb)
    # the copies really evolved apart (different actions)
    p = benv.fluid.pressure.view(B, -1)
    assert (p[0] - p[1]).abs().max().item() > 1e3


class OracleEnv:
    """ReservoirEnv (rl_env.py:90-154) restated on the FP64 oracle with the exact
    pressure solve: actions move the oracle wells' controls by the same rule
    (rl_env.py:129-140), the observation is (p / 1e7, Sw) in float32
    (rl_env.py:123-127) and the reward the oil mass produced in the step / 1e6
    (rl_env.py:142-154: porosity * So * rho_o(p) * cell volume summed)."""

    def __init__(self, cfg):
        from oracle import impes as oi

        self.oi = oi
        pvt = oi.PVT(cfg["fluid"]["pvt_table"]) if cfg["fluid"].get("pvt_table") else None
        self.case = oi.Case(cfg, pvt=pvt)
        self.sim = oi.Simulation(self.case)
        self.dt = float(cfg["simulation"].get("time_step_days", 1.0)) * 86400.0
        self.mass = self.oil_mass()

    def oil_mass(self):
        st = self.sim.state
        so = 1.0 - st["sw"] - st["sg"]
        return float(np.sum(self.case.pv * so * self.case.densities(st["p"])[1]))

    def obs(self):
        st = self.sim.state
        return np.concatenate([st["p"].ravel() / 1e7, st["sw"].ravel()]).astype(np.float32)

    def step(self, action):
        for a, w in zip(action, self.case.wells):
            if w.control_type == "rate":
                w.control_value = float(w.control_value + w.control_value * 0.2 * float(a))
            elif w.control_type == "bhp":
                w.control_value = float(w.control_value + 5.0 * float(a))
        self.sim.step(self.dt)
        m = self.oil_mass()
        r = (self.mass - m) / 1e6
        self.mass = m
        return self.obs(), r


def test_env_trajectory_matches_oracle(tmp_path, dev):
    """A ReservoirEnv trajectory under a fixed random action sequence against the
    oracle-stepped environment with the same actions (each side propagates its own
    state): observations, rewards, well controls and substep counts at every step."""
    from rl_env import ReservoirEnv

    path, cfg = _cfg_path(tmp_path)
    steps = 6
    actions = np.random.default_rng(11).uniform(-1.0, 1.0, size=(steps, N_WELLS))
    env = ReservoirEnv(path, max_steps=steps, device=dev)
    obs, _ = env.reset(seed=0)
    orc = OracleEnv(cfg)
    assert np.array_equal(obs, orc.obs())
    for t in range(steps):
        o, r, term, trunc, info = env.step(actions[t])
        oo, ro = orc.step(actions[t])
        assert info["converged"]
        assert [w.control_value for w in env.well_manager.wells] == [w.control_value for w in orc.case.wells]
        assert env.simulator.last_substeps == orc.sim.history[-1]["N"], t
        assert np.allclose(o, oo, rtol=0, atol=2e-6), (t, np.max(np.abs(o - oo)))
        assert r == pytest.approx(ro, rel=1e-6, abs=1e-12), (t, r, ro)
    assert term
