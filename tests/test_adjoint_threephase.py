"""The three-phase adjoint (adjoint/threephase.py) on the CPU with a direct FP64
solve.

Pins: the torch forward equals the oracle's step (oracle/impes.py) on a
config-2-like case (PVT table, surface-rate gas injectors and oil producers, a
BHP producer, gravity); the FaceSolve reverse (transposed solve + operator
cotangents) equals torch autograd through a dense differentiable solve; the
gradient matches central finite differences along random unit directions of
ln k (h = 1e-4) and along every well control; checkpointed and plain runs give
the same gradient. Rates are scaled down and the initial pressure lowered so
that no choke binds and the pressures stay inside the PVT table: there the
objective is smooth and the finite differences are meaningful."""
import numpy as np
import pytest
import torch

from adjoint import threephase as tp
from adjoint.twophase import dense_solve, scipy_solver
from oracle import impes as oi
from simulator.cases import baseline_case

STEPS = 3


@pytest.fixture(scope="module")
def cfg_fields():
    cfg, f = baseline_case(2, dims=(10, 8, 5))
    for w in cfg["wells"]:
        if w.get("control_type") == "rate":
            w["control_value"] *= 0.02
    cfg["fluid"]["pressure"] = 15.0
    return cfg, f


@pytest.fixture(scope="module")
def case(cfg_fields):
    return tp.ThreePhaseCase.from_config(*cfg_fields, device="cpu")


def test_forward_equals_oracle(cfg_fields, case):
    cfg, f = cfg_fields
    oc = oi.Case(cfg, f, pvt=oi.PVT(cfg["fluid"]["pvt_table"]))
    st = oc.initial_state()
    p, sw, sg = case.initial_state()
    for _ in range(STEPS):
        ps = oi.pressure_system(oc, st, case.dt)
        x = oi.solve_direct(oi.system_csr(oc, ps), ps["b"])
        sat = oi.saturation_step(oc, st, x.reshape(oc.dims), case.dt, ps)
        st = {"p": x.reshape(oc.dims), "sw": sat["sw"], "sg": sat["sg"]}
        with torch.no_grad():
            p, sw, sg, N = tp.step(case, case.lnk0, case.rates0, p, sw, sg, scipy_solver())
        assert N == sat["N"]
        assert np.abs(p.numpy() - st["p"]).max() <= 1e-12 * np.abs(st["p"]).max()
        assert np.abs(sw.numpy() - st["sw"]).max() <= 1e-9
        assert np.abs(sg.numpy() - st["sg"]).max() <= 1e-9


def test_face_solve_reverse_equals_dense_autograd(case):
    J1, gl1, gv1, c1 = tp.gradient(case, case.lnk0, case.rates0, scipy_solver(), STEPS, checkpoint_every=None)
    J2, gl2, gv2, c2 = tp.gradient(case, case.lnk0, case.rates0, dense_solve, STEPS, checkpoint_every=None,
                                   direct=True)
    assert c1 == c2
    assert abs(J1 - J2) <= 1e-12 * abs(J2)
    assert (gl1 - gl2).norm() <= 1e-6 * gl2.norm()
    assert (gv1 - gv2).norm() <= 1e-6 * gv2.norm()


def test_checkpointed_gradient_equals_plain(case):
    _, gl1, gv1, c1 = tp.gradient(case, case.lnk0, case.rates0, scipy_solver(), STEPS, checkpoint_every=None)
    _, gl2, gv2, c2 = tp.gradient(case, case.lnk0, case.rates0, scipy_solver(), STEPS, checkpoint_every=2)
    assert c1 == c2
    assert (gl1 - gl2).norm() <= 1e-12 * gl1.norm()
    assert (gv1 - gv2).norm() <= 1e-12 * gv1.norm()


def _J(case, lnk, values, branches=None):
    with torch.no_grad():
        p, sw, sg, counts = tp.simulate(case, lnk, values, scipy_solver(), STEPS, branches=branches)
        return tp.oil_in_place(case, p, sw, sg).item(), counts


def test_gradient_matches_finite_differences(case):
    J, gl, gv, counts = tp.gradient(case, case.lnk0, case.rates0, scipy_solver(), STEPS, checkpoint_every=2)
    assert gl.norm() > 1e-4 * abs(J)  # a sensitive objective (no binding choke or table clamp everywhere)
    rng = np.random.default_rng(3)
    h = 1e-4
    for _ in range(4):
        v = torch.tensor(rng.standard_normal(case.dims))
        v /= v.norm()
        bp, bm = [], []
        jp, cp = _J(case, case.lnk0 + h * v, case.rates0, bp)
        jm, cm = _J(case, case.lnk0 - h * v, case.rates0, bm)
        # no substep count crossed; clamp decisions may flip where a clamp is only
        # marginally active (a kink too small to show at this tolerance)
        assert cp == cm == counts and len(bp) == len(bm) > 0
        fd = (jp - jm) / (2 * h)
        assert abs(fd - float((gl * v).sum())) <= 1e-4 * abs(fd)
    for wn in range(len(case.wells)):
        hv = 1e-4 * max(1.0, abs(float(case.rates0[wn])))
        e = torch.zeros_like(case.rates0)
        e[wn] = hv
        fd = (_J(case, case.lnk0, case.rates0 + e)[0] - _J(case, case.lnk0, case.rates0 - e)[0]) / (2 * hv)
        assert abs(fd - float(gv[wn])) <= 1e-5 * max(abs(fd), 1e-3 * float(gv.abs().max())), wn


def test_replayed_branches_give_the_adjoint_derivative(case):
    """With the base run's decisions replayed (tp.BranchLog), the +-h runs stay on
    the smooth branch the adjoint differentiates: along the gradient with h = 0.1,
    where evaluated decisions flip and move the difference by tens of percent,
    the replayed difference is off only by its O(h^2) truncation; at h = 1e-3 it
    equals the adjoint derivative to 1e-6. Replay at the base point reproduces
    the base run and its branch record."""
    J, gl, gv, counts = tp.gradient(case, case.lnk0, case.rates0, scipy_solver(), STEPS, checkpoint_every=None)
    base = tp.BranchLog()
    J0, c0 = _J(case, case.lnk0, case.rates0, base)
    again = tp.BranchLog(replay=base)
    J1, c1 = _J(case, case.lnk0, case.rates0, again)
    assert J1 == J0 and c1 == c0 and tp.decisions_changed(list(again), list(base)) == 0
    v = gl / gl.norm()
    an = float(gl.norm())
    for h, tol in ((1e-1, 2e-2), (1e-3, 1e-6)):
        jp, _ = _J(case, case.lnk0 + h * v, case.rates0, tp.BranchLog(replay=base))
        jm, _ = _J(case, case.lnk0 - h * v, case.rates0, tp.BranchLog(replay=base))
        assert abs((jp - jm) / (2 * h) - an) <= tol * an, h
