"""The hand-written HIP adjoint of the three-phase IMPES run (adjoint/device.py,
csrc/adjoint.hip) on small history-match problems, objective: the producer-rate
misfit against synthetic data (2% noise) from a truth ln k, gradient at a prior.

Problems:
  gas_injection    the gas-injection benchmark case at 10x8x5 cells, rates x0.02,
                   p0 = 15 MPa (PVT, Stone II, capillarity, gravity, surface-rate
                   gas injectors, oil-rate producers, a BHP producer)
  producer_limits  the same with Sw = 0.3, Sg = 0.05 and 60x60x4 m cells: the oil
                   producers flow and their gas/liquid rate limits bind
  gas3             the reference's gas_injection_3phase configuration (no PVT
                   table, o-w capillarity, gravity, rate injectors, BHP producers)
  pcog             Stone II with gas-oil capillarity (golden case pcog)

Checks per problem: device observations = the torch restatement's (1e-10); the
device gradient = torch autograd through the restatement with an FP64 direct
solve on the host (objective 1e-10, ln k and well-control gradients 1e-9
relative); central differences of the restatement with the base run's
decisions replayed along 20 random unit directions of ln k (h = 1e-4) within
1e-4 of the adjoint; on gas_injection also FP32 = FP64 hierarchy (1e-8).

producer_limits amplifies the forward's pressure-solve roundoff ~30-fold per
step (device and restatement observations differ by 2e-14, 2e-13, 7e-12, 2e-10
relative over the four steps: growth, not a formula difference, which would
show at the first step), so its two runs are compared at 1e-8 (observations)
and 1e-6 (gradients); the finite-difference check holds it to the same 1e-4, at
h = 1e-5: kinks the replay does not fix (relative-permeability clamps at the
saturation bounds, PVT table nodes) lie within 1e-4 of the prior in ln k
(replayed differences along its directions: 0.6 / 3e-2 / 1e-5 relative error at
h = 1e-3 / 1e-4 / 1e-5)."""
import json
import os

import numpy as np
import pytest
import torch

from adjoint import device as ad
from adjoint import threephase as tp
from adjoint.twophase import scipy_solver
from simulator.cases import SEEDS

pytestmark = pytest.mark.gpu

STEPS = 4
TIGHT = {"rtol": 1e-13, "adj_rtol": 1e-13, "etol": 0.0, "maxit": 600}
# (observations, gradients) relative bars against the restatement
BARS = {"producer_limits": (1e-8, 1e-6)}
FD_H = {"producer_limits": 1e-5}
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    meta = json.load(open(os.path.join(GOLD, "steps.json")))[name]
    cfg = json.loads(json.dumps(meta["config"]))
    if meta["pvt"]:
        cfg["fluid"]["pvt_table"] = meta["pvt"]
        cfg["fluid"].pop("pvt_path", None)
    cfg["simulation"]["time_step_days"] = meta["dt_days"]
    return cfg


def make(name):
    """(config, fields, ln k truth, ln k prior) of a problem."""
    if name in ("gas_injection", "producer_limits"):
        cfg, fields, truth, prior = ad.history_match_case(dims=(10, 8, 5), steps=STEPS, rate_factor=0.02,
                                                          p0_mpa=15.0)
        if name == "producer_limits":
            cfg["fluid"].update({"s_w": 0.3, "s_g": 0.05})
            cfg["reservoir"]["grid_size"] = [60.0, 60.0, 4.0]
            for i, w in enumerate(cfg["wells"]):
                if w["name"].startswith("OPROD"):
                    w["limits"] = {"wgpr": 640.0} if i % 2 == 0 else {"wlpr": 0.5}
        return cfg, fields, truth, prior
    cfg = golden(name)
    case = tp.ThreePhaseCase.from_config(cfg, None, device="cpu")
    truth = case.lnk0.numpy()
    prior = truth + 0.2 * np.random.default_rng(SEEDS["prior"]).standard_normal(truth.shape)
    return cfg, None, truth, prior


@pytest.fixture(scope="module", params=["gas_injection", "producer_limits", "gas3", "pcog"])
def problem(request):
    cfg, fields, lnk_truth, lnk_prior = make(request.param)
    dev = ad.DeviceAdjoint(cfg, fields, solver=dict(TIGHT, precision=32))
    truth = torch.tensor(np.asarray(lnk_truth).reshape(-1), dtype=torch.float64, device="cuda")
    obs_truth = dev.observe(truth, dev.values0, STEPS)
    data, sigma = ad.rate_data(obs_truth, SEEDS["noise"])
    case = tp.ThreePhaseCase.from_config(cfg, fields, device="cpu")
    prior = torch.tensor(np.asarray(lnk_prior).reshape(case.dims), dtype=torch.float64)
    return {"name": request.param, "cfg": cfg, "fields": fields, "dev": dev, "data": data, "sigma": sigma,
            "prior": prior, "case": case}


def torch_misfit(P, lnk, branches=None):
    case = P["case"]
    with torch.no_grad():
        d = tp.observe(case, lnk, case.rates0, scipy_solver(), STEPS, branches=branches)
    return float(tp.rate_misfit(d, torch.tensor(P["data"]), torch.tensor(P["sigma"])))


def test_observations_equal_restatement(problem):
    P = problem
    dev, case = P["dev"], P["case"]
    d_dev = dev.observe(P["prior"].reshape(-1).cuda(), dev.values0, STEPS)
    with torch.no_grad():
        d_t = tp.observe(case, P["prior"], case.rates0, scipy_solver(), STEPS).numpy()
    assert np.abs(d_t).max() > 0
    assert np.abs(d_dev - d_t).max() <= BARS.get(P["name"], (1e-10, 1e-9))[0] * np.abs(d_t).max()


def test_device_gradient_equals_autograd(problem):
    P = problem
    dev, case = P["dev"], P["case"]
    J, g_lnk, g_val, _ = dev.gradient(P["prior"].reshape(-1).cuda(), dev.values0, STEPS, P["data"], P["sigma"],
                                      checkpoint_every=2)
    Jt, gl_t, gv_t = tp.misfit_gradient(case, P["prior"], case.rates0, scipy_solver(), STEPS,
                                        torch.tensor(P["data"]), torch.tensor(P["sigma"]), checkpoint_every=None)
    obs_bar, grad_bar = BARS.get(P["name"], (1e-10, 1e-9))
    assert abs(J - float(Jt)) <= obs_bar * abs(float(Jt))
    gl = g_lnk.cpu()
    assert gl_t.norm() > 0
    err_l = float((gl - gl_t).norm() / gl_t.norm())
    gv = torch.tensor(g_val)
    err_v = float((gv - gv_t).norm() / gv_t.norm())
    print(f"{P['name']}: ln k gradient rel err {err_l:.2e}, controls {err_v:.2e}")
    P["grad"] = (J, gl, gv)
    assert err_l <= grad_bar and err_v <= grad_bar


def test_device_gradient_matches_replayed_finite_differences(problem):
    P = problem
    if "grad" not in P:
        pytest.skip("needs test_device_gradient_equals_autograd")
    J, gl, _ = P["grad"]
    base = tp.BranchLog()
    J0 = torch_misfit(P, P["prior"], base)
    assert abs(J0 - J) <= BARS.get(P["name"], (1e-10, 1e-9))[0] * abs(J)
    rng = np.random.default_rng(SEEDS["fd"])
    h = FD_H.get(P["name"], 1e-4)
    worst = 0.0
    for _ in range(20):
        v = torch.tensor(rng.standard_normal(tuple(P["prior"].shape)))
        v /= v.norm()
        jp = torch_misfit(P, P["prior"] + h * v, tp.BranchLog(replay=base))
        jm = torch_misfit(P, P["prior"] - h * v, tp.BranchLog(replay=base))
        fd = (jp - jm) / (2 * h)
        an = float((gl * v).sum())
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-300))
    print(f"{P['name']}: worst replayed FD rel err {worst:.2e}")
    assert worst < 1e-4, worst


def test_fp32_and_fp64_hierarchies_agree(problem):
    P = problem
    if P["name"] != "gas_injection":
        pytest.skip("one problem suffices")
    dev64 = ad.DeviceAdjoint(P["cfg"], P["fields"], solver=dict(TIGHT, precision=64))
    lnk = P["prior"].reshape(-1).cuda()
    J32, g32, v32, _ = P["dev"].gradient(lnk, P["dev"].values0, STEPS, P["data"], P["sigma"], checkpoint_every=3)
    J64, g64, v64, _ = dev64.gradient(lnk, dev64.values0, STEPS, P["data"], P["sigma"], checkpoint_every=3)
    assert abs(J32 - J64) <= 1e-10 * abs(J64)
    assert (g32 - g64).norm() <= 1e-8 * g64.norm()
