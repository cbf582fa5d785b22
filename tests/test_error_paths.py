"""Error behaviour of the drop-in API and of the C ABI, without a GPU.

The reference raises ValueError on inconsistent PVT tables (pvt.py:43-46), on
wells without coordinates or control (well.py:82,92) and on initial saturations
above one (fluid.py:84-85); the mirrors keep those. The device path has no CPU
fallback: the simulator and the AMG refuse a CPU device. At the C ABI every
entry point returns a negative status with a message instead of failing inside
a kernel; the calls below are rejected by argument checks that run before any
HIP call. The last test covers the host logic of the pressure solve's start."""
import ctypes as C

import pytest

from simulator.fluid import Fluid
from simulator.pvt import PVTTable
from simulator.reservoir import Reservoir
from simulator.well import WellManager

PVT = {"pressure_MPa": [10.0, 20.0, 30.0], "Bo": [1.2, 1.19, 1.18], "Bw": [1.01, 1.0, 0.99],
       "Bg": [0.01, 0.005, 0.003], "mu_o_cP": [1.0, 1.1, 1.2], "mu_w_cP": [0.5, 0.5, 0.5],
       "mu_g_cP": [0.02, 0.02, 0.02], "Rs_m3m3": [50.0, 80.0, 100.0]}


def test_pvt_table_rejects_inconsistent_columns():
    PVTTable(data=PVT)
    bad = dict(PVT, pressure_MPa=[10.0, 30.0, 20.0])
    with pytest.raises(ValueError):
        PVTTable(data=bad)
    with pytest.raises(ValueError):
        PVTTable(data=dict(PVT, Bo=[1.2, 1.19]))
    with pytest.raises(ValueError):
        PVTTable(data={k: v[:1] for k, v in PVT.items()})


def _reservoir():
    return Reservoir({"dimensions": [4, 3, 2], "grid_size": [10.0, 10.0, 5.0]}, device="cpu")


def test_well_manager_rejects_missing_coordinates_or_control():
    res = _reservoir()
    good = {"name": "P1", "type": "producer", "coordinates": [1, 1, 0], "radius": 0.1,
            "control_type": "bhp", "control_value": 15.0}
    WellManager([good], res)
    no_pos = {k: v for k, v in good.items() if k != "coordinates"}
    with pytest.raises(ValueError):
        WellManager([no_pos], res)
    no_ctl = {k: v for k, v in good.items() if k not in ("control_type", "control_value")}
    with pytest.raises(ValueError):
        WellManager([no_ctl], res)


def test_fluid_rejects_saturations_above_one():
    res = _reservoir()
    Fluid({"pressure": 20.0, "s_w": 0.3, "s_g": 0.2}, res, device="cpu")
    with pytest.raises(ValueError):
        Fluid({"pressure": 20.0, "s_w": 0.7, "s_g": 0.4}, res, device="cpu")


def test_device_path_refuses_a_cpu_device():
    import torch

    from simulator.simulation import Simulator
    from solver.classical_amg import ClassicalAMG

    res = _reservoir()
    fluid = Fluid({"pressure": 20.0}, res, device="cpu")
    wells = WellManager([], res)
    with pytest.raises(RuntimeError):
        Simulator(res, fluid, wells, {"solver_type": "impes"}, device="cpu")
    A = torch.eye(4, dtype=torch.float64).to_sparse_csr()
    with pytest.raises(RuntimeError):
        ClassicalAMG(A, device="cpu")


def test_c_abi_rejects_bad_arguments_before_any_device_call():
    import impes_hip

    lib = impes_hip.load()
    h = C.c_void_p()
    assert lib.impes_amg_create(None, C.byref(h)) == 0
    try:
        dummy = C.c_void_p(16)  # never dereferenced: the checks below fail first
        assert lib.impes_amg_vcycle(h, 64, dummy, dummy, 1, None) < 0
        assert b"not set up" in lib.impes_last_error()
        assert lib.impes_amg_vcycle_transposed(h, 64, dummy, dummy, 1, None) < 0
        assert b"not set up" in lib.impes_last_error()
        assert lib.impes_amg_setup_csr(h, -1, 0, dummy, None, None, None) < 0
        assert b"negative size" in lib.impes_last_error()
        assert lib.impes_amg_setup_csr(h, 4, 8, None, None, None, None) < 0
        assert b"null CSR pointer" in lib.impes_last_error()
        assert lib.impes_amg_vcycle(None, 64, dummy, dummy, 1, None) < 0
    finally:
        assert lib.impes_amg_destroy(h) == 0


def test_pressure_initial_guess_modes():
    """The FGMRES start of the simulator's pressure solve (host logic only)."""
    import torch

    from simulator.simulation import Simulator

    sim = Simulator.__new__(Simulator)
    sim._p_hist = None
    p = torch.tensor([10.0, 20.0], dtype=torch.float64)
    assert sim._initial_guess(p, 2.0, {}) is p  # no accepted step yet
    sim._p_hist = (torch.tensor([8.0, 21.0], dtype=torch.float64), 4.0)
    # p + (p - p_last) * dt / dt_last with dt / dt_last = 1/2
    assert torch.equal(sim._initial_guess(p, 2.0, {}), torch.tensor([11.0, 19.5], dtype=torch.float64))
    # after a shorter step (one the attempt loop cut) the current pressure
    sim._p_hist = (torch.tensor([8.0, 21.0], dtype=torch.float64), 0.5)
    assert sim._initial_guess(p, 2.0, {}) is p
    assert sim._initial_guess(p, 2.0, {"initial_guess": "previous"}) is p
    assert torch.equal(sim._initial_guess(p, 2.0, {"initial_guess": "zero"}), torch.zeros(2, dtype=torch.float64))
    sim._p_hist = (torch.zeros(3, dtype=torch.float64), 1.0)  # another grid: no extrapolation
    assert sim._initial_guess(p, 2.0, {}) is p
