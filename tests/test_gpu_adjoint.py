"""The adjoint on the GPU: its forward (torch device ops + the HIP AMG/FGMRES
solve) is the drop-in Simulator's step, and its gradient with the device solver
equals the CPU direct-solve adjoint and central finite differences."""
import numpy as np
import pytest
import torch

from adjoint import twophase as tp
from adjoint.history import problem

pytestmark = pytest.mark.gpu


def test_adjoint_forward_is_the_simulator_step(dev):
    from simulator.cases import baseline_case, build

    dims, steps = (16, 16, 4), 4
    cfg, f = baseline_case(3, dims=dims, steps=steps)
    cfg["simulation"]["amg"] = {"tol": 1e-12, "precision": "fp32", "max_iter": 500}
    sim = build(cfg, f, "cuda")
    case = tp.TwoPhaseCase.from_config(cfg, f, device=dev)
    lnk = torch.tensor(np.log(f["perm_x"]), device=dev)
    rates = torch.tensor([w["rate"] for w in case.wells], dtype=torch.float64, device=dev)
    solve = tp.device_solver(rtol=1e-12)
    for _ in range(steps):
        p0, sw0 = sim.fluid.pressure.double().clone(), sim.fluid.s_w.double().clone()
        assert sim.run_step(case.dt)
        p, sw = tp.step(case, lnk, rates, p0, sw0, solve)
        dp = float((p - sim.fluid.pressure).abs().max() / sim.fluid.pressure.abs().max())
        ds = float((sw - sim.fluid.s_w).abs().max())
        assert dp < 1e-8 and ds < 1e-6, (dp, ds)


def test_device_adjoint_matches_cpu_adjoint_and_fd(dev):
    P = problem(dims=(20, 20, 4), steps=8)
    _, gl_c, gr_c = tp.gradient(P["case"], P["lnk_prior"], P["rates"], P["data"], P["sigma"], tp.scipy_solver(),
                                P["steps"], checkpoint_every=3)
    Pd = problem(dims=(20, 20, 4), steps=8, device=dev, solve=tp.device_solver(rtol=1e-13))
    solve = tp.device_solver(rtol=1e-13)
    case, lnk, rates, data, sigma = Pd["case"], Pd["lnk_prior"], Pd["rates"], Pd["data"], Pd["sigma"]
    J, gl, gr = tp.gradient(case, lnk, rates, data, sigma, solve, Pd["steps"], checkpoint_every=3)
    assert float((gl.cpu() - gl_c).norm() / gl_c.norm()) < 1e-6
    assert float((gr.cpu() - gr_c).norm() / gr_c.norm()) < 1e-6

    def Jf(l):
        o, _ = tp.simulate(case, l, rates, solve, Pd["steps"])
        return float(tp.misfit(o, data, sigma))

    rng = np.random.default_rng(3)
    for _ in range(5):
        v = torch.tensor(rng.standard_normal(tuple(lnk.shape)), device=dev)
        v /= v.norm()
        fd = (Jf(lnk + 1e-4 * v) - Jf(lnk - 1e-4 * v)) / 2e-4
        an = float((gl * v).sum())
        assert abs(fd - an) <= 1e-4 * abs(fd), (fd, an)
