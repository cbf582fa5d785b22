"""The adjoint history-match problem: two-phase oil-water on a 64x64x8 grid
(fluids of the 60x60 waterflood case), truth and prior log-permeability fields
from different seeds, 2 water injectors and 3 liquid producers at reservoir
rates, 30 steps of 5 days. Data are the producers' water cuts and the injectors'
bottom-hole pressures simulated from the truth with 1% relative Gaussian noise."""
from __future__ import annotations

import numpy as np
import torch

from simulator.cases import SEEDS, baseline_case, gaussian_field

from .twophase import TwoPhaseCase, simulate

SIGMA_FLOOR = {"producer": 1e-3, "injector": 1e3}  # water cut [-], BHP [Pa]


def problem(dims=None, steps=None, device="cpu", solve=None):
    """Returns dict(case, lnk_truth, lnk_prior, rates, data, sigma, steps, obs_truth)."""
    cfg, fields = baseline_case(3, dims=dims, steps=steps)
    dims = tuple(cfg["reservoir"]["dimensions"])
    steps = int(cfg["simulation"]["steps"])
    case = TwoPhaseCase.from_config(cfg, fields, device=device)
    lnk_truth = torch.as_tensor(np.log(fields["perm_x"]), dtype=torch.float64, device=device)
    prior = np.log(100.0) + 1.0 * gaussian_field(dims, 4.0, SEEDS["prior"], axes=(0, 1, 2))
    lnk_prior = torch.as_tensor(prior, dtype=torch.float64, device=device)
    rates = torch.tensor([w["rate"] for w in case.wells], dtype=torch.float64, device=device)
    if solve is None:
        from .twophase import scipy_solver

        solve = scipy_solver()
    with torch.no_grad():
        obs, _ = simulate(case, lnk_truth, rates, solve, steps)
    rng = np.random.default_rng(SEEDS["noise"])
    noise = torch.as_tensor(rng.standard_normal(tuple(obs.shape)), dtype=torch.float64, device=device)
    data = obs * (1.0 + 0.01 * noise)
    floor = torch.tensor([SIGMA_FLOOR[w["type"]] for w in case.wells], dtype=torch.float64, device=device)
    sigma = torch.maximum(0.01 * data.abs(), floor.expand_as(data))
    return {"case": case, "lnk_truth": lnk_truth, "lnk_prior": lnk_prior, "rates": rates, "data": data,
            "sigma": sigma, "steps": steps, "obs_truth": obs}
