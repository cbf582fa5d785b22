"""Pressure solver facade (reference: src/solver/pressure_amg.py:35-160).

``amg_solve(A, b, ...)`` keeps the reference signature, hierarchy cache and
return convention (x on b's device and dtype). The default method is FP64
restarted FGMRES on the row-scaled system D^-1 A x = D^-1 b, right
preconditioned by one RS-AMG V-cycle (FP32 hierarchy unless precision=64),
stopping on ||D^-1 (b - A x)|| <= tol * ||D^-1 b||; ``method="vcycle"``
restores the reference's stationary V-cycle iteration (relative residual of
the equilibrated system). ``amg_solve.last_stats`` holds the solver report.
``hierarchy="reference"`` builds the reference's hierarchy instead of this
package's default (see solver.classical_amg). ``error_tol`` > 0 adds the
FGMRES pressure-error check (estimated relative error, infinity norm).
"""
from __future__ import annotations

from typing import Any, Dict, Optional, Tuple

import torch

from .classical_amg import ClassicalAMG, _device, to_device_csr

_AMG_CACHE: Dict[Tuple[Any, ...], Any] = {}
_SOLVER = {}


def _entry(A, theta, max_levels, coarsest_size, dev, hierarchy="default"):
    from impes_hip import AmgHierarchy

    if hierarchy not in ("default", "reference"):
        raise ValueError(f"hierarchy must be 'default' or 'reference', got {hierarchy!r}")
    rp, ci, v, shape = to_device_csr(A, dev)
    key = (tuple(shape), float(theta), int(max_levels), int(coarsest_size), str(dev), hierarchy)
    ent = _AMG_CACHE.get(key)
    if ent is not None and torch.equal(ent["rp"], rp) and torch.equal(ent["ci"], ci):
        ent["amg"].update_csr(v)
        ent["hit"] = True
    else:
        extra = dict(AmgHierarchy.REFERENCE) if hierarchy == "reference" else {}
        h = AmgHierarchy(theta=theta, max_levels=max_levels, coarsest_size=coarsest_size, **extra)
        h.setup_csr(rp, ci, v)
        ent = {"amg": h, "rp": rp, "ci": ci, "hit": False}
        _AMG_CACHE[key] = ent
    ent["v"] = v
    return ent


def amg_solve(
    A: torch.Tensor,
    b: torch.Tensor,
    *,
    tol: float = 1e-6,
    max_cycles: int = 20,
    theta: float = 0.25,
    max_levels: int = 10,
    coarsest_size: int = 200,
    device: Optional[str] = "auto",
    enable_equilibration: Optional[bool] = None,
    method: str = "fgmres",
    precision: int = 32,
    restart: int = 30,
    maxiter: Optional[int] = None,
    hierarchy: str = "default",
    error_tol: float = 0.0,
) -> torch.Tensor:
    """Solve A x = b; returns x on b.device with b.dtype."""
    from impes_hip import FgmresSolver, csr_diagonal

    dev = _device(device)
    ent = _entry(A, theta, max_levels, coarsest_size, dev, hierarchy)
    h = ent["amg"]
    b64 = b.to(dev, dtype=torch.float64).contiguous().reshape(-1)
    if method == "vcycle":
        x, rel, cyc = h.cycles(b64, None, min_cycles=1, tol=tol, max_cycles=max_cycles, precision=64)
        amg_solve.last_stats = {"method": "vcycle", "cycles": cyc, "rel_residual": rel, "converged": rel < tol}
    else:
        solver = _SOLVER.setdefault(str(dev), FgmresSolver())
        dinv = 1.0 / csr_diagonal(ent["rp"], ent["ci"], ent["v"])
        x = torch.zeros_like(b64)
        st = solver.solve_csr(h, ent["rp"], ent["ci"], ent["v"], dinv, b64, x, precision=precision, rtol=tol,
                              restart=restart, maxit=maxiter if maxiter is not None else max(int(max_cycles), 300),
                              etol=error_tol)
        st["method"] = "fgmres"
        amg_solve.last_stats = st
    return x.to(b.device, dtype=b.dtype)


amg_solve.last_stats = {}


def amg_preconditioner(
    A: torch.Tensor,
    *,
    theta: float = 0.25,
    max_levels: int = 10,
    coarsest_size: int = 200,
    device: Optional[str] = "auto",
    hierarchy: str = "default",
) -> Dict[str, Any]:
    """pressure_amg.py:92-160: right preconditioner r -> z ~ A^-1 r (one V-cycle through
    ClassicalAMG.apply), with the reference's theta back-off on a quality test."""
    dev = _device(device)
    rp, ci, v, shape = to_device_csr(A, dev)
    A_csr = torch.sparse_csr_tensor(rp.long(), ci.long(), v, size=shape)

    def assess(amg: ClassicalAMG, cycles: int = 2) -> float:
        from impes_hip import spmv_csr

        n = shape[0]
        tests = [torch.ones(n, dtype=torch.float64, device=dev), torch.randn(n, dtype=torch.float64, device=dev),
                 torch.linspace(0, 1, n, dtype=torch.float64, device=dev)]
        worst = 0.0
        for t in tests:
            t = t / (t.norm() + 1e-30)
            z = amg.apply(t, cycles=cycles)
            Az = spmv_csr(rp, ci, v, z)
            worst = max(worst, float((Az - t).norm() / (t.norm() + 1e-30)))
        return worst

    candidates = [theta] + ([0.2] if theta > 0.2 else []) + [0.15, 0.1]
    best, best_score, chosen = None, float("inf"), None
    for th in candidates:
        amg = ClassicalAMG(A_csr, theta=th, max_levels=max_levels, coarsest_size=coarsest_size, device=dev,
                           hierarchy=hierarchy)
        score = assess(amg)
        if score < best_score:
            best, best_score = amg, score
        if score <= 0.05:
            chosen = amg
            break
    if chosen is None:
        raise RuntimeError(f"AMG preconditioner quality too low (best score={best_score:.3e})")

    def _apply(r: torch.Tensor) -> torch.Tensor:
        z = chosen.apply(r.to(dev, dtype=torch.float64), cycles=1)
        return z.to(r.device, dtype=r.dtype)

    return {"apply": _apply, "device": chosen.device}
