"""The hand-written adjoint of the two-phase IMPES step (CPU, FP64, direct solves).

Pins: the torch forward restatement equals the oracle's step (1e-12); the
adjoint gradient equals torch autograd through the same forward with a dense
direct solve (1e-9 relative); central finite differences along 20 random unit
directions of ln k agree to 1e-4 relative. The FD step is h = 1e-4: the
discrete forward carries ~1e-12 relative roundoff in J (direct solve, clamped
substeps), so h = 1e-6 is roundoff-limited (~1e-4 error) while h = 1e-4 leaves
a truncation error ~1e-7."""
import numpy as np
import pytest
import torch

from adjoint import twophase as tp
from adjoint.history import problem
from oracle import impes as oi
from simulator.cases import baseline_case


@pytest.fixture(scope="module")
def small():
    return problem(dims=(8, 8, 3), steps=6)


def test_forward_matches_oracle():
    dims, steps = (10, 10, 3), 6
    cfg, f = baseline_case(3, dims=dims, steps=steps)
    case = tp.TwoPhaseCase.from_config(cfg, f)
    lnk = torch.tensor(np.log(f["perm_x"]))
    rates = torch.tensor([w["rate"] for w in case.wells], dtype=torch.float64)
    solve = tp.scipy_solver()
    orc = oi.Simulation(oi.Case(cfg, f))
    p = torch.full(dims, case.p0, dtype=torch.float64)
    sw = torch.full(dims, case.sw0, dtype=torch.float64)
    substeps = []
    for _ in range(steps):
        rec = orc.step(case.dt)
        tape = tp.StepTape()
        p, sw = tp.step(case, lnk, rates, p, sw, solve, tape)
        assert tape.N == rec["N"]
        substeps.append(tape.N)
        assert np.max(np.abs(p.numpy() - rec["p"])) / np.max(np.abs(rec["p"])) < 1e-12
        assert np.max(np.abs(sw.numpy() - rec["sat"]["sw"])) < 1e-12
        orc.state = {"p": p.numpy().copy(), "sw": sw.numpy().copy(), "sg": np.zeros(dims)}
    assert max(substeps) > 1  # the clamped-substep branch is exercised


def test_adjoint_matches_autograd(small):
    P = small
    case, lnk, rates = P["case"], P["lnk_prior"], P["rates"]
    J, gl, gr = tp.gradient(case, lnk, rates, P["data"], P["sigma"], tp.scipy_solver(), P["steps"],
                            checkpoint_every=2)
    lv, rv = lnk.clone().requires_grad_(True), rates.clone().requires_grad_(True)
    obs, _ = tp.simulate(case, lv, rv, tp.dense_solve, P["steps"])
    Ja = tp.misfit(obs, P["data"], P["sigma"])
    gla, gra = torch.autograd.grad(Ja, (lv, rv))
    assert abs(float(J - Ja)) <= 1e-10 * abs(float(Ja))
    assert float((gl - gla).norm() / gla.norm()) < 1e-9
    assert float((gr - gra).norm() / gra.norm()) < 1e-9


def test_adjoint_matches_finite_differences(small):
    P = small
    case, lnk, rates, data, sigma, steps = P["case"], P["lnk_prior"], P["rates"], P["data"], P["sigma"], P["steps"]
    solve = tp.scipy_solver()
    _, gl, gr = tp.gradient(case, lnk, rates, data, sigma, solve, steps, checkpoint_every=3)

    def J(l, r):
        o, _ = tp.simulate(case, l, r, solve, steps)
        return float(tp.misfit(o, data, sigma))

    rng = np.random.default_rng(3)
    h = 1e-4
    for _ in range(20):
        v = torch.tensor(rng.standard_normal(tuple(lnk.shape)))
        v /= v.norm()
        fd = (J(lnk + h * v, rates) - J(lnk - h * v, rates)) / (2 * h)
        an = float((gl * v).sum())
        assert abs(fd - an) <= 1e-4 * abs(fd), (fd, an)
    for w in range(len(rates)):
        e = torch.zeros_like(rates)
        e[w] = 1.0
        fd = (J(lnk, rates + 1e-3 * e) - J(lnk, rates - 1e-3 * e)) / 2e-3
        assert abs(fd - float(gr[w])) <= 1e-6 * abs(fd), (w, fd, float(gr[w]))


def test_checkpoint_interval_does_not_change_the_gradient(small):
    P = small
    args = (P["case"], P["lnk_prior"], P["rates"], P["data"], P["sigma"], tp.scipy_solver(), P["steps"])
    ref = tp.gradient(*args, checkpoint_every=None)
    for k in (1, 2, 4):
        out = tp.gradient(*args, checkpoint_every=k)
        assert torch.equal(out[1], ref[1]) and torch.equal(out[2], ref[2])


def test_autograd_function_returns_the_adjoint_gradient(small):
    P = small
    lv, rv = P["lnk_prior"].clone().requires_grad_(True), P["rates"].clone().requires_grad_(True)
    J = tp.HistoryMatchMisfit.apply(lv, rv, P["case"], P["data"], P["sigma"], tp.scipy_solver(), P["steps"], 2)
    (3.0 * J).backward()
    _, gl, gr = tp.gradient(P["case"], P["lnk_prior"], P["rates"], P["data"], P["sigma"], tp.scipy_solver(),
                            P["steps"], checkpoint_every=2)
    assert torch.allclose(lv.grad, 3.0 * gl, rtol=1e-14, atol=0)
    assert torch.allclose(rv.grad, 3.0 * gr, rtol=1e-14, atol=0)


def test_truth_has_zero_noise_free_misfit():
    P = problem(dims=(6, 6, 2), steps=3)
    obs, _ = tp.simulate(P["case"], P["lnk_truth"], P["rates"], tp.scipy_solver(), P["steps"])
    assert torch.equal(obs, P["obs_truth"])
    assert float(tp.misfit(obs, P["obs_truth"], P["sigma"])) == 0.0
