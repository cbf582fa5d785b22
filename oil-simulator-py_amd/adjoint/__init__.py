"""Discrete adjoints of the IMPES step: hand-written reverse sweep for two-phase
oil-water cases (``twophase``) and the three-phase black-oil step with torch
autograd around the HIP pressure solve (``threephase``)."""
from . import threephase  # noqa: F401
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
