"""Discrete adjoints of the IMPES step: hand-written reverse sweep for two-phase
oil-water cases (``twophase``), the three-phase black-oil step on the HIP
kernels with the hand-written reverse-sweep kernels of csrc/adjoint.hip
(``device``), and its torch restatement differentiated by autograd
(``threephase``, the reference the device adjoint is checked against)."""
from . import device, threephase  # noqa: F401
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
