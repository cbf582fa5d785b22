"""Thin Python layer over the gfx950 HIP library ``libimpes_hip.so``."""
from ._lib import (  # noqa: F401
    LIB_PATH,
    AmgParams,
    FluidStruct,
    Grid,
    WellStruct,
    ImpesError,
    SolveStats,
    call,
    exported_symbols,
    load,
    ptr,
    stream_ptr,
)
from .ops import *  # noqa: F401,F403
