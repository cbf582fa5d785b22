"""GPU IMPES step (HIP kernels through the drop-in Simulator) against the FP64
oracle with an exact pressure solve.

Bars (per step, from the same input state): pressure 1e-8 relative, Sw/Sg 1e-6
absolute, identical substep counts, ledgers 1e-6 relative; fluid properties
and the assembled system to 1e-12 relative."""
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import impes as oi

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLD, "steps.json")))
FL = np.load(os.path.join(GOLD, "fluid.npz"))
FLCFG = json.load(open(os.path.join(GOLD, "fluid.json")))
PVT_SYNTH = os.path.join(GOLD, "pvt_synthetic.json")


def golden_config(name):
    cfg = json.loads(json.dumps(META[name]["config"]))
    if META[name]["pvt"]:
        cfg["fluid"]["pvt_table"] = META[name]["pvt"]
        cfg["fluid"].pop("pvt_path", None)
    return cfg, META[name]["dt_days"] * 86400.0


def oracle_for(cfg, fields=None):
    pvt = oi.PVT(cfg["fluid"]["pvt_table"]) if cfg["fluid"].get("pvt_table") else None
    return oi.Simulation(oi.Case(cfg, fields, pvt=pvt))


def gpu_sim(cfg, fields=None, **sim_over):
    from simulator import cases

    cfg = json.loads(json.dumps(cfg, default=float)) if fields is None else dict(cfg)
    cfg["simulation"] = dict(cfg.get("simulation", {}), **sim_over)
    return cases.build(cfg, fields, "cuda")


def np_state(sim):
    f = sim.fluid
    return {"p": f.pressure.double().cpu().numpy(), "sw": f.s_w.double().cpu().numpy(),
            "sg": f.s_g.double().cpu().numpy()}


@pytest.mark.parametrize("variant", ["corey", "stone2", "pvt", "nopvt"])
def test_fluid_properties_match_oracle(dev, variant):
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir

    fc = dict({"pressure": 20.0, "s_w": 0.2, "s_g": 0.0}, **FLCFG[variant])
    if "pvt_path" in fc:
        fc["pvt_path"] = PVT_SYNTH
    res = Reservoir({"dimensions": [41, 41, 1], "grid_size": [10, 10, 10]}, dev)
    fl = Fluid(fc, res, dev)
    case = oi.Case({"reservoir": {"dimensions": [41, 41, 1], "grid_size": [10, 10, 10]}, "fluid": fc, "wells": []})
    SW, SG, P = (FL[k].astype(np.float64)[..., None] for k in ("SW", "SG", "P"))
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    kro, krw, krg = fl.get_rel_perms(t(SW), t(SG))
    okro, okrw, okrg = case.relperms(SW, SG)
    for g, o in ((kro, okro), (krw, okrw), (krg, okrg)):
        assert np.allclose(g.cpu().numpy(), o, rtol=1e-13, atol=1e-15)
    assert np.allclose(fl.get_capillary_pressure(t(SW)).cpu().numpy(), case.pc_ow(SW), rtol=1e-13)
    assert np.allclose(fl.get_capillary_pressure_og(t(SG)).cpu().numpy(), case.pc_og(SG), rtol=1e-13)
    ct = fl.calc_total_compressibility(t(P), t(SW), t(SG))
    assert np.allclose(torch.as_tensor(ct).cpu().numpy(), case.ct(P, SW, SG), rtol=1e-12, atol=1e-25)
    for name, o in zip(("rho_w", "rho_o", "rho_g"), case.densities(P)):
        g = getattr(fl, f"calc_{ {'rho_w': 'water', 'rho_o': 'oil', 'rho_g': 'gas'}[name]}_density")(t(P))
        assert np.allclose(g.cpu().numpy(), o, rtol=1e-14), name
    if case.pvt is not None:
        for key in ("Bo", "Bw", "Bg", "mu_o", "mu_w", "mu_g", "Rs", "Rv"):
            assert np.allclose(fl._eval_pvt(t(P), key).cpu().numpy(), case.pvt(key, P), rtol=1e-15, atol=0)


@pytest.mark.parametrize("name", sorted(META))
def test_pressure_system_matches_oracle(dev, name):
    cfg, dt = golden_config(name)
    sim = gpu_sim(cfg)
    orc = oracle_for(cfg)
    for _ in range(2):
        st = np_state(sim)
        ps = oi.pressure_system(orc.case, st, dt)
        sim._impes_pressure_step(dt)
        g = {k: v.cpu().numpy() for k, v in sim.pressure_system().items()}
        for k, o in (("diag", ps["diag"]), ("ox", ps["ox"]), ("oy", ps["oy"]), ("oz", ps["oz"]), ("rhs", ps["b"])):
            scale = np.abs(o) + (np.abs(ps["acc"] * st["p"].ravel()) if k == "rhs" else 0)
            assert np.all(np.abs(g[k] - o) <= 1e-12 * scale + 1e-300), (name, k)
        sim.run_step(dt)


def solver_cfg(precision=32, **over):
    """The benchmark's pressure-solver settings (FP32 or FP64 hierarchy)."""
    from simulator.cases import BENCH_AMG

    return dict(BENCH_AMG, precision="fp64" if precision == 64 else "fp32", **over)


def close(a, b, rel=1e-6, scale=0.0):
    return math.isclose(a, b, rel_tol=rel, abs_tol=1e-9 * (1.0 + abs(b)) + rel * scale)


def run_parity(cfg, fields, dt, steps, amg, p_tol=1e-8, s_tol=1e-6, whole_step=True):
    """Steps the device simulator with the solver settings `amg`; before each step
    the oracle takes the simulator's state.

    Pressure: against the oracle's exact solve (p_tol, relative). Transport:
    against the oracle's transport step evaluated at the simulator's own new
    pressure: saturations (s_tol, absolute), substep count, component and mass
    ledger increments (simulation.py:1396-1529) and every well's diagnostics
    (last_mode, last_q_total, last_surface_rates; well.py:53-55). Whole step
    (whole_step=True): saturations also against the oracle step with the exact
    pressure, which holds where the reference's transport is continuous in the
    pressure at solver-error level. At the end: the cumulative volume balance."""
    sim = gpu_sim(cfg, fields, pressure_solver="amg", amg=amg)
    orc = oracle_for(cfg, fields)
    worst = {"p": 0.0, "s": 0.0, "s_whole": 0.0}
    prev_cb = {ph: dict(v) for ph, v in sim.component_balance.items()}
    prev_mb = {ph: dict(v) for ph, v in sim.mass_balance.items()}
    wells = sim.well_manager.get_wells()
    for s in range(steps):
        before = np_state(sim)
        orc.state = before
        rec = orc.step(dt)
        assert sim.run_step(dt)
        st = np_state(sim)
        po = rec["p"]
        dp = np.max(np.abs(st["p"] - po)) / np.max(np.abs(po))
        worst["p"] = max(worst["p"], dp)
        assert dp < p_tol, (s, dp, sim.last_solve)
        sat = oi.saturation_step(orc.case, before, st["p"], dt, rec["ps"])
        ds = max(np.max(np.abs(st["sw"] - sat["sw"])), np.max(np.abs(st["sg"] - sat["sg"])))
        worst["s"] = max(worst["s"], ds)
        assert ds < s_tol, (s, ds)
        ds_whole = max(np.max(np.abs(st["sw"] - rec["sat"]["sw"])), np.max(np.abs(st["sg"] - rec["sat"]["sg"])))
        worst["s_whole"] = max(worst["s_whole"], ds_whole)
        for o, check in ((sat, True), (rec["sat"], whole_step)):
            # N = ceil(max rat); when max rat sits on an integer (a choked well filling
            # its window exactly) the two roundings may land either side of it
            r = o["rat_max"]
            on_int = abs(r - round(r)) <= 1e-9 * max(1.0, abs(r))
            assert not check or sim.last_substeps == o["N"] or (on_int and abs(sim.last_substeps - o["N"]) <= 1), (
                s, sim.last_substeps, o["N"], r)
        if whole_step:
            assert ds_whole < s_tol, (s, ds_whole)
        cb = sim.component_balance
        for ph in ("oil", "gas"):
            for io in ("in", "out"):
                inc = cb[ph][io] - prev_cb[ph][io]
                ref = sat["comp"][f"{ph}_{io}"]
                assert close(inc, ref), (s, ph, io, inc, ref)
        prev_cb = {ph: dict(v) for ph, v in cb.items()}
        mb = sim.mass_balance
        mass = sat["mass"]
        for ph in ("water", "oil", "gas"):
            # the ledgers sum clipped per-cell increments; a cell whose increment sits
            # at rounding level may land on either side of zero: scale by the total
            scale = 1e-3 * float(mass[ph + "_in"] + mass[ph + "_out"])
            for io in ("in", "out"):
                inc = mb[ph][io] - prev_mb[ph][io]
                ref = float(mass[f"{ph}_{io}"])
                assert close(inc, ref, scale=scale), (s, ph, io, inc, ref)
        prev_mb = {ph: dict(v) for ph, v in mb.items()}
        for w, wd in zip(wells, sat["wells"]):
            if wd["last_mode"] is not None:
                assert w.last_mode == wd["last_mode"], (s, w.name, w.last_mode, wd["last_mode"])
            if wd["last_q_total"] is not None:
                assert close(w.last_q_total, wd["last_q_total"]), (s, w.name, w.last_q_total, wd["last_q_total"])
                for ph, v in wd["last_surface_rates"].items():
                    assert close(w.last_surface_rates[ph], v, scale=abs(wd["last_q_total"])), (s, w.name, ph)
    err = sim.volume_balance_error()
    assert err["water"] < 1e-10 and err["gas"] < 1e-10, err
    return sim, worst


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", sorted(META))
def test_reference_cases_match_oracle(dev, name, precision):
    cfg, dt = golden_config(name)
    run_parity(cfg, None, dt, steps=3, amg=solver_cfg(precision))


@pytest.mark.parametrize("index,dims,steps", [(0, (60, 60, 1), 8), (1, (20, 20, 6), 6), (2, (16, 16, 10), 4),
                                              (3, (24, 24, 4), 4), (1, (40, 40, 10), 4), (2, (32, 32, 24), 3)])
def test_benchmark_cases_match_oracle(dev, index, dims, steps):
    """The benchmark cases (shrunk) with exactly the benchmark's solver settings
    (FP32 V-cycles inside FP64 FGMRES to a 1e-8 relative residual, extrapolated
    start, pressure-error check at 1e-8). The residual alone would not do: on the
    config-1 cases a 1e-8 residual in the row-scaled norm leaves pressure errors
    up to ~5e-7 relative (cells below the bubble point carry a negative total
    compressibility), which the error check catches and removes."""
    from simulator.cases import baseline_case

    cfg, fields = baseline_case(index, dims=dims)
    dt = cfg["simulation"]["time_step_days"] * 86400.0
    amg = solver_cfg(32)
    assert cfg["simulation"]["amg"] == (amg if index else solver_cfg(64))
    # config 2's transport amplifies pressure errors ~1e6-fold next to the gas
    # injectors (a 1e-12 relative pressure change moves saturations by ~5e-6:
    # tests/test_oracle_impes.py::test_gas_injection_transport_amplifies_pressure_error),
    # so its transport is compared at the device pressure only
    sim, _ = run_parity(cfg, fields, dt, steps, amg=amg, whole_step=index != 2)
    assert all(r["solve"]["checks"] >= 1 for r in sim.step_log)


def test_residual_stop_alone_misses_pressure_bar(dev):
    """Without the error check (error_tol 0) the 1e-8 residual stop leaves the
    shrunk config-1 pressure above the 1e-8 bar within a few steps: the check is
    what holds the bar, not slack in the test."""
    from simulator.cases import baseline_case

    cfg, fields = baseline_case(1, dims=(20, 20, 6))
    dt = cfg["simulation"]["time_step_days"] * 86400.0
    with pytest.raises(AssertionError):
        run_parity(cfg, fields, dt, 6, amg=solver_cfg(32, error_tol=0.0))


@pytest.mark.parametrize("limit,phase", [("wopr", "oil"), ("wlpr", "water"), ("wgpr", "gas"), ("liqr", "liquid")])
def test_producer_rate_limits_match_oracle(dev, limit, phase):
    """Surface-rate producer cut back by each rate limit (simulation.py:1262-1320),
    three-phase with PVT: the device step against the oracle, wells included."""
    cfg = {"simulation": {"solver_type": "impes"},
           "reservoir": {"dimensions": [5, 5, 1], "grid_size": [10, 10, 10], "porosity": 0.25,
                         "permeability": 200.0, "k_vertical_fraction": 1.0},
           "fluid": {"pressure": 25.0, "s_w": 0.45, "s_g": 0.15, "pvt_table": json.load(open(PVT_SYNTH))},
           "wells": [{"name": "P", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1,
                      "control_type": "rate", "control_value": 8.0, "rate_type": "surface",
                      "surface_phase": "liquid", "limits": {limit: 1.0 if phase != "gas" else 40.0}},
                     {"name": "I", "type": "injector", "i": 0, "j": 0, "k": 0, "radius": 0.1,
                      "control_type": "rate", "control_value": 5.0, "injected_phase": "water"}]}
    sim, _ = run_parity(cfg, None, 86400.0, 2, amg=solver_cfg(64))
    w = sim.well_manager.get_wells()[0]
    assert w.last_surface_rates[phase] == pytest.approx(1.0 if phase != "gas" else 40.0, rel=1e-9)


def test_extrapolated_start_after_a_cut_step(dev):
    """After a step shorter than the next one (as the attempt loop leaves behind
    when it cuts dt), the extrapolated start falls back to the current pressure:
    from identical states the next full step then solves exactly as with
    initial_guess "previous"."""
    from simulator.cases import baseline_case

    cfg, fields = baseline_case(1, dims=(20, 20, 6))
    dt = cfg["simulation"]["time_step_days"] * 86400.0
    out = {}
    for mode in ("extrapolate", "previous"):
        sim = gpu_sim(cfg, fields, pressure_solver="amg", amg=solver_cfg(32, initial_guess="previous"))
        assert sim.run_step(dt) and sim.run_step(dt / 8)
        sim._amg_cfg["initial_guess"] = mode
        assert sim.run_step(dt)
        out[mode] = (sim.last_solve["iterations"], sim.fluid.pressure.clone())
    assert out["extrapolate"][0] == out["previous"][0]
    assert torch.equal(out["extrapolate"][1], out["previous"][1])


def test_reference_cg_solver_path(dev):
    """pressure_solver 'cg' (the reference default): Jacobi-PCG with the absolute stop."""
    cfg, dt = golden_config("ow_2d")
    sim = gpu_sim(cfg, pressure_solver="cg", cg_tolerance=1e-9, cg_max_iter=2000)
    orc = oracle_for(cfg)
    orc.state = np_state(sim)
    rec = orc.step(dt)
    assert sim.run_step(dt)
    assert sim.last_solve["converged"]
    p = sim.fluid.pressure.cpu().numpy()
    assert np.max(np.abs(p - rec["p"])) / np.max(np.abs(rec["p"])) < 1e-7


def test_cg_nonconvergence_triggers_dt_retry(dev):
    """simulation.py:833-865: an unconverged solve halves dt (forced with cg_max_iter=1)."""
    cfg, dt = golden_config("ow_2d")
    sim = gpu_sim(cfg, pressure_solver="cg", cg_tolerance=1e-30, cg_max_iter=1, max_time_step_attempts=3)
    assert sim.run_step(dt) is False
