"""Reinforcement-learning environments over the IMPES step (reference:
src/rl_env.py:15-154).

``ReservoirEnv`` keeps the reference's gymnasium-style API: constructor
(config_path, max_steps, device, render_mode), ``reset`` -> (obs, info),
``step(action)`` -> (obs, reward, terminated, truncated, info), the action
scaling (rate wells +-20 % of the current control per unit action, BHP wells
+-5 MPa, applied cumulatively as the reference does), the observation
(pressure / 1e7 and Sw, float32, C order) and the reward (oil mass produced in
the step / 1e6). The step itself is the device IMPES step.

``BatchedReservoirEnv`` (not in the reference) steps ``num_envs`` copies of the
reservoir in ONE device step: the copies are stacked along x into one grid
whose faces between copies carry no flow, so assembly, the AMG-preconditioned
pressure solve and the transport run once for all of them, and every copy keeps
its own substep count. Each copy follows the trajectory a ``ReservoirEnv``
would with the same actions, up to the pressure-solve tolerance (the stacked
solve meets the relative residual over all copies together). The copies share
one time step: a failed pressure solve ends the step of all of them with the
reference's penalty.

gymnasium is used when installed; otherwise a minimal stand-in of ``Env`` and
``spaces.Box`` (the parts this module uses) takes its place.
"""
from __future__ import annotations

import copy
import json
from typing import Dict, Tuple

import numpy as np
import torch

from simulator.fluid import Fluid
from simulator.reservoir import Reservoir
from simulator.simulation import Simulator
from simulator.well import WellManager

try:  # gymnasium is optional
    import gymnasium as gym
    from gymnasium import spaces

    _EnvBase = gym.Env
except ImportError:  # pragma: no cover - exercised where gymnasium is absent
    gym = None

    class _Box:
        """The subset of gymnasium.spaces.Box used here."""

        def __init__(self, low, high, shape, dtype=np.float32):
            self.shape = tuple(shape)
            self.dtype = np.dtype(dtype)
            self.low = np.full(self.shape, low, dtype=self.dtype)
            self.high = np.full(self.shape, high, dtype=self.dtype)

        def sample(self, rng=None):
            rng = rng or np.random.default_rng()
            lo = np.where(np.isfinite(self.low), self.low, -1.0)
            hi = np.where(np.isfinite(self.high), self.high, 1.0)
            return rng.uniform(lo, hi).astype(self.dtype)

        def contains(self, x) -> bool:
            x = np.asarray(x)
            return x.shape == self.shape and bool(np.all(x >= self.low) and np.all(x <= self.high))

    class spaces:  # noqa: N801 - mirrors the gymnasium module name
        Box = _Box

    class _EnvBase:
        metadata: dict = {}

        def reset(self, *, seed=None, options=None):
            self.np_random = np.random.default_rng(seed)


def _load(config_path: str) -> dict:
    with open(config_path, "r", encoding="utf-8") as f:
        return json.load(f)


def _apply_action_to_wells(action, wells) -> None:
    """rl_env.py:132-144: rate wells move by up to +-20 % of their current control,
    BHP wells by up to +-5 MPa; other control types are left alone."""
    for a, well in zip(action, wells):
        if well.control_type == "rate":
            base_rate = well.control_value
            well.control_value = float(base_rate + base_rate * 0.2 * float(a))
        elif well.control_type == "bhp":
            well.control_value = float(well.control_value + 5.0 * float(a))


class ReservoirEnv(_EnvBase):
    """rl_env.py:15-154 on the device IMPES step."""

    metadata = {"render_modes": ["human"], "render_fps": 4}

    def __init__(self, config_path: str, max_steps: int = 365, device: torch.device | None = None,
                 render_mode: str | None = None) -> None:
        super().__init__()
        self.config_path = config_path
        self.max_steps = max_steps
        self.render_mode = render_mode
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self._build(_load(config_path))
        n_wells = len(self.well_manager.wells)
        self.action_space = spaces.Box(low=-1.0, high=1.0, shape=(n_wells,), dtype=np.float32)
        nx, ny, nz = self.reservoir.dimensions
        self.observation_space = spaces.Box(-np.inf, np.inf, shape=(nx * ny * nz * 2,), dtype=np.float32)
        self._step_count = 0
        self._prev_oil_mass = self._calc_oil_mass().item()

    def _build(self, cfg: dict) -> None:
        self.reservoir = Reservoir(cfg["reservoir"], device=self.device)
        self.well_manager = WellManager(cfg["wells"], self.reservoir)
        self.fluid = Fluid(cfg["fluid"], self.reservoir, device=self.device)
        self.simulator = Simulator(reservoir=self.reservoir, fluid=self.fluid, well_manager=self.well_manager,
                                   sim_params=cfg.get("simulation", {}), device=self.device)

    def reset(self, *, seed: int | None = None, options: Dict | None = None):
        super().reset(seed=seed)
        self._build(_load(self.config_path))
        self._step_count = 0
        self._prev_oil_mass = self._calc_oil_mass().item()
        return self._get_observation(), {}

    def step(self, action: np.ndarray) -> Tuple[np.ndarray, float, bool, bool, Dict]:
        self._apply_action(action)
        dt_sec = self.simulator.sim_params.get("time_step_days", 1.0) * 24 * 3600
        if not self.simulator.run_step(dt_sec):
            return self._get_observation(), -1e3, True, False, {"converged": False}
        self._step_count += 1
        obs = self._get_observation()
        reward = self._calc_reward()
        terminated = self._step_count >= self.max_steps
        return obs, reward, terminated, False, {"converged": True}

    def render(self, mode: str = "human"):
        if mode != "human":
            raise NotImplementedError
        p_mean = float(self.fluid.pressure.mean() / 1e6)
        print(f"Mean pressure: {p_mean:.2f} MPa, step {self._step_count}")

    def close(self):
        pass

    def _get_observation(self) -> np.ndarray:
        p = self.fluid.pressure.cpu().numpy().flatten() / 1e7
        sw = self.fluid.s_w.cpu().numpy().flatten()
        return np.concatenate([p, sw]).astype(np.float32)

    def _apply_action(self, action: np.ndarray):
        _apply_action_to_wells(action, self.well_manager.wells)
        self.simulator.update_wells()

    def _calc_oil_mass(self) -> torch.Tensor:
        return torch.sum(self.reservoir.porosity * self.fluid.s_o * self.fluid.rho_o * self.reservoir.cell_volume)

    def _calc_reward(self) -> float:
        oil_mass = self._calc_oil_mass().item()
        produced_oil = self._prev_oil_mass - oil_mass
        self._prev_oil_mass = oil_mass
        return produced_oil / 1e6


def stack_config(cfg: dict, num_envs: int) -> dict:
    """The configuration of num_envs copies of cfg's reservoir stacked along x: the
    reservoir's x dimension times num_envs, and every well repeated per copy with
    its x index shifted by copy * nx and its name suffixed with '#<copy>'."""
    if num_envs < 1:
        raise ValueError("num_envs must be >= 1")
    out = copy.deepcopy(cfg)
    nx, ny, nz = (int(v) for v in cfg["reservoir"].get("dimensions", [10, 10, 1]))
    out["reservoir"]["dimensions"] = [nx * num_envs, ny, nz]
    wells = []
    for e in range(num_envs):
        for w in cfg["wells"]:
            v = copy.deepcopy(w)
            if "coordinates" in v:
                i, j, k = v["coordinates"]
                v["coordinates"] = [int(i) + e * nx, j, k]
            elif "i" in v:
                v["i"] = int(v["i"]) + e * nx
            else:
                raise ValueError(f"invalid well coordinates: {w}")
            v["name"] = f"{w['name']}#{e}"
            wells.append(v)
    out["wells"] = wells
    return out


class BatchedReservoirEnv(_EnvBase):
    """num_envs copies of ReservoirEnv stepped together in one device step.

    Actions have shape (num_envs, n_wells); observations (num_envs, 2 * cells);
    rewards, terminated and truncated (num_envs,). ``info["substeps"]`` holds each
    copy's transport substep count of the step."""

    metadata = {"render_modes": ["human"], "render_fps": 4}

    def __init__(self, config_path: str, num_envs: int, max_steps: int = 365, device: torch.device | None = None):
        super().__init__()
        self.config_path = config_path
        self.num_envs = int(num_envs)
        self.max_steps = max_steps
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        cfg = _load(config_path)
        self.env_dims = tuple(int(v) for v in cfg["reservoir"].get("dimensions", [10, 10, 1]))
        self.wells_per_env = len(cfg["wells"])
        self._build(cfg)
        nx, ny, nz = self.env_dims
        self.single_action_space = spaces.Box(low=-1.0, high=1.0, shape=(self.wells_per_env,), dtype=np.float32)
        self.action_space = spaces.Box(low=-1.0, high=1.0, shape=(self.num_envs, self.wells_per_env),
                                       dtype=np.float32)
        self.single_observation_space = spaces.Box(-np.inf, np.inf, shape=(nx * ny * nz * 2,), dtype=np.float32)
        self.observation_space = spaces.Box(-np.inf, np.inf, shape=(self.num_envs, nx * ny * nz * 2),
                                            dtype=np.float32)
        self._step_count = 0
        self._prev_oil_mass = self._calc_oil_mass()

    def _build(self, cfg: dict) -> None:
        stacked = stack_config(cfg, self.num_envs)
        self.reservoir = Reservoir(stacked["reservoir"], device=self.device)
        self.well_manager = WellManager(stacked["wells"], self.reservoir)
        self.fluid = Fluid(stacked["fluid"], self.reservoir, device=self.device)
        self.simulator = Simulator(reservoir=self.reservoir, fluid=self.fluid, well_manager=self.well_manager,
                                   sim_params=stacked.get("simulation", {}), device=self.device,
                                   env_nx=self.env_dims[0])

    def reset(self, *, seed: int | None = None, options: Dict | None = None):
        super().reset(seed=seed)
        self._build(_load(self.config_path))
        self._step_count = 0
        self._prev_oil_mass = self._calc_oil_mass()
        return self._get_observation(), {}

    def step(self, actions: np.ndarray):
        actions = np.asarray(actions, dtype=np.float64).reshape(self.num_envs, self.wells_per_env)
        wells = self.well_manager.wells
        for e in range(self.num_envs):
            _apply_action_to_wells(actions[e], wells[e * self.wells_per_env:(e + 1) * self.wells_per_env])
        self.simulator.update_wells()
        dt_sec = self.simulator.sim_params.get("time_step_days", 1.0) * 24 * 3600
        B = self.num_envs
        if not self.simulator.run_step(dt_sec):
            return (self._get_observation(), np.full(B, -1e3), np.ones(B, dtype=bool), np.zeros(B, dtype=bool),
                    {"converged": False})
        self._step_count += 1
        obs = self._get_observation()
        oil_mass = self._calc_oil_mass()
        rewards = (self._prev_oil_mass - oil_mass) / 1e6
        self._prev_oil_mass = oil_mass
        terminated = np.full(B, self._step_count >= self.max_steps)
        info = {"converged": True, "substeps": list(self.simulator.last_substeps_env or [self.simulator.last_substeps])}
        return obs, rewards, terminated, np.zeros(B, dtype=bool), info

    def render(self, mode: str = "human"):
        if mode != "human":
            raise NotImplementedError
        p = self.fluid.pressure.view(self.num_envs, -1).mean(dim=1) / 1e6
        print("Mean pressure per copy (MPa): " + ", ".join(f"{float(v):.2f}" for v in p) +
              f", step {self._step_count}")

    def close(self):
        pass

    def _get_observation(self) -> np.ndarray:
        B = self.num_envs
        p = self.fluid.pressure.reshape(B, -1).cpu().numpy() / 1e7
        sw = self.fluid.s_w.reshape(B, -1).cpu().numpy()
        return np.concatenate([p, sw], axis=1).astype(np.float32)

    def _calc_oil_mass(self) -> np.ndarray:
        m = self.reservoir.porosity * self.fluid.s_o * self.fluid.rho_o * self.reservoir.cell_volume
        return m.reshape(self.num_envs, -1).sum(dim=1).cpu().numpy()
