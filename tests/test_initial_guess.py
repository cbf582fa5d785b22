"""The FGMRES starting pressure of the simulator (host logic, CPU tensors): the
modes "zero", "previous", "extrapolate" and "extrapolate2" and their fallbacks.
The solve itself runs to the same stopping test whatever the start."""
import torch

from simulator.simulation import Simulator


def bare(hist=None, hist2=None):
    sim = Simulator.__new__(Simulator)  # the method reads only the pressure history
    sim._p_hist = hist
    sim._p_hist2 = hist2
    return sim


def series(n, trend=0.5, alt=0.3):
    """Start pressures of n steps: a linear trend plus an alternating component."""
    base = torch.linspace(0.0, 1.0, 7, dtype=torch.float64)
    return [base + trend * k + alt * (-1.0) ** k for k in range(n)]


def test_zero_and_previous():
    p = torch.ones(5, dtype=torch.float64)
    sim = bare((torch.zeros(5, dtype=torch.float64), 1.0))
    assert torch.equal(sim._initial_guess(p, 1.0, {"initial_guess": "zero"}), torch.zeros(5, dtype=torch.float64))
    assert sim._initial_guess(p, 1.0, {"initial_guess": "previous"}) is p


def test_extrapolate_is_linear_in_time():
    p0, p1 = series(2)
    sim = bare((p0, 2.0))
    got = sim._initial_guess(p1, 1.0, {"initial_guess": "extrapolate"})
    assert torch.allclose(got, p1 + (p1 - p0) * 0.5, rtol=0, atol=1e-15)
    # after a shorter step (cut by the attempt loop): the current pressure
    assert bare((p0, 0.5))._initial_guess(p1, 1.0, {"initial_guess": "extrapolate"}) is p1
    # no history yet
    assert bare()._initial_guess(p1, 1.0, {}) is p1


def test_extrapolate2_is_exact_for_trend_plus_alternation():
    p0, p1, p2, p3 = series(4)
    sim = bare((p1, 1.0), (p0, 1.0))
    got = sim._initial_guess(p2, 1.0, {"initial_guess": "extrapolate2"})
    assert torch.allclose(got, p3, rtol=0, atol=1e-14)
    # the linear extrapolation triples the alternating part instead
    lin = sim._initial_guess(p2, 1.0, {"initial_guess": "extrapolate"})
    assert torch.allclose(lin - p3, torch.full_like(p3, 4 * 0.3 * (-1.0) ** 2), rtol=0, atol=1e-14)


def test_extrapolate2_falls_back_to_linear():
    p0, p1, p2 = series(3)
    # only one step of history
    got = bare((p1, 1.0))._initial_guess(p2, 1.0, {"initial_guess": "extrapolate2"})
    assert torch.allclose(got, p2 + (p2 - p1), rtol=0, atol=1e-15)
    # unequal steps behind it
    got = bare((p1, 1.0), (p0, 2.0))._initial_guess(p2, 1.0, {"initial_guess": "extrapolate2"})
    assert torch.allclose(got, p2 + (p2 - p1), rtol=0, atol=1e-15)
