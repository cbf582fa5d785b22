"""Hand-written discrete adjoint of the IMPES step (two-phase oil-water)."""
from .twophase import (  # noqa: F401
    HistoryMatchMisfit,
    TwoPhaseCase,
    dense_solve,
    device_solver,
    gradient,
    misfit,
    observe,
    scipy_solver,
    simulate,
    step,
    step_vjp,
)
