"""The three-phase adjoint with the HIP solver: FP32 RS-AMG inside FP64 FGMRES
(rtol 1e-12) for the forward and the transposed solves against the FP64 direct
solve on the host, on a config-2-like case (PVT, surface-rate wells, gravity).
Bars: objective 1e-10 relative, gradients 1e-6 relative, substep counts exact."""
import pytest
import torch

from adjoint import threephase as tp
from adjoint.twophase import device_solver, scipy_solver
from simulator.cases import baseline_case

pytestmark = pytest.mark.gpu


def test_device_gradient_equals_direct(dev):
    cfg, f = baseline_case(2, dims=(16, 12, 6))
    for w in cfg["wells"]:
        if w.get("control_type") == "rate":
            w["control_value"] *= 0.02
    cfg["fluid"]["pressure"] = 15.0
    cg = tp.ThreePhaseCase.from_config(cfg, f, device=dev)
    cc = tp.ThreePhaseCase.from_config(cfg, f, device="cpu")
    J1, gl1, gv1, c1 = tp.gradient(cg, cg.lnk0, cg.rates0, device_solver(rtol=1e-12, precision=32), 3,
                                   checkpoint_every=2)
    J2, gl2, gv2, c2 = tp.gradient(cc, cc.lnk0, cc.rates0, scipy_solver(), 3, checkpoint_every=None)
    assert c1 == c2
    assert abs(J1.item() - J2.item()) <= 1e-10 * abs(J2.item())
    assert gl2.norm() > 1e-4 * abs(J2.item())
    assert (gl1.cpu() - gl2).norm() <= 1e-6 * gl2.norm()
    assert (gv1.cpu() - gv2).norm() <= 1e-6 * gv2.norm()
