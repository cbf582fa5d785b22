"""The reference's own test suite, run against the drop-in API on the GPU.

Each test restates one test of the reference (file:line cited) with the same
configuration and the same assertion, constructing Reservoir / WellManager /
Fluid / Simulator exactly as the reference test does but with device "cuda".
The reference reads configs/pvt/pvt_synthetic.json and
configs/gas_injection_3phase.json; the same data come from the fixture copy of
the PVT table (tests/golden/pvt_synthetic.json) and from the golden metadata
(case "gas3")."""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PVT = os.path.join(HERE, "golden", "pvt_synthetic.json")
META = json.load(open(os.path.join(HERE, "golden", "steps.json")))


def mk(cfg, sim_over=None):
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir
    from simulator.simulation import Simulator
    from simulator.well import WellManager

    dev = torch.device("cuda")
    res = Reservoir(cfg["reservoir"], dev)
    wells = WellManager(cfg["wells"], res)
    fl = Fluid(cfg["fluid"], res, dev)
    sim = Simulator(res, fl, wells, dict(cfg["simulation"], **(sim_over or {})), dev)
    return res, wells, fl, sim


def gas3_config():
    cfg = json.loads(json.dumps(META["gas3"]["config"]))
    if META["gas3"]["pvt"]:
        cfg["fluid"]["pvt_table"] = META["gas3"]["pvt"]
        cfg["fluid"].pop("pvt_path", None)
    return cfg


def masses(res, fl):
    V = res.cell_volume
    mw = torch.sum(res.porosity * fl.s_w * fl.rho_w * V).item()
    mo = torch.sum(res.porosity * fl.s_o * fl.rho_o * V).item()
    mg = torch.sum(res.porosity * fl.s_g * fl.rho_g * V).item()
    return mw + mo + mg


def test_amg_matches_cg_pressure_field():
    """tests/test_amg_pressure_solver.py:16-58: one step with 'cg' and with 'gmres_amg'."""
    out = {}
    for name in ("cg", "gmres_amg"):
        cfg = gas3_config()
        sim_cfg = dict(cfg.get("simulation", {}) or {})
        sim_cfg.update({"solver_type": "impes", "total_time_days": 1.0, "time_step_days": 1.0,
                        "pressure_solver": name})
        if name != "cg":
            sim_cfg["amg"] = {"tol": 1e-6, "max_cycles": 20, "theta": 0.25, "max_levels": 10,
                              "coarsest_size": 200, "device": "auto"}
        cfg["simulation"] = sim_cfg
        _, _, fl, sim = mk(cfg)
        assert sim.run_step(dt=86400.0)
        out[name] = fl.pressure.detach().cpu().numpy().reshape(-1)
    rel = np.linalg.norm(out["gmres_amg"] - out["cg"]) / max(np.linalg.norm(out["cg"]), 1e-12)
    assert rel < 5e-4


@pytest.mark.parametrize("hierarchy", ["default", "reference"])
def test_classical_amg_quality_on_laplacian(hierarchy):
    """tests/test_amg_pressure_solver.py:89-106 through the ClassicalAMG facade,
    with this package's hierarchy and with the reference's."""
    from solver.classical_amg import ClassicalAMG, _apply_reference_fix

    n = 12
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n):
            p, d = i * n + j, 0.0
            for ii, jj in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                if 0 <= ii < n and 0 <= jj < n:
                    rows.append(p); cols.append(ii * n + jj); vals.append(-1.0); d += 1.0
            rows.append(p); cols.append(p); vals.append(d)
    A = torch.sparse_coo_tensor(torch.tensor([rows, cols]), torch.tensor(vals, dtype=torch.float64),
                                (n * n, n * n)).coalesce().to_sparse_csr()
    A_fixed = _apply_reference_fix(A, anchor_idx=0)
    amg = ClassicalAMG(A_fixed, theta=0.25, max_levels=10, coarsest_size=40, anchor_idx=None, hierarchy=hierarchy)
    N = A_fixed.size(0)
    vs = [torch.ones(N, dtype=torch.float64), torch.linspace(0, 1, N, dtype=torch.float64),
          torch.randn(N, dtype=torch.float64, generator=torch.Generator().manual_seed(0))]
    Ad = A_fixed.to_dense().to(amg.device)
    for v in vs:
        v = (v / (v.norm() + 1e-30)).to(amg.device)
        z = amg.apply(v, cycles=3)
        rel = (Ad @ z - v).norm() / (v.norm() + 1e-30)
        assert rel < 5e-2, float(rel)


def test_amg_solve_and_preconditioner_facades():
    """pressure_amg.amg_solve / amg_preconditioner (reference pressure_amg.py:35-160)
    on an IMPES-shaped system: converged FP64 solution, cached hierarchy reuse,
    result on b's device and dtype."""
    import scipy.sparse.linalg as spl

    from matrices import impes_like
    from solver.pressure_amg import amg_preconditioner, amg_solve

    A, *_ = impes_like(20, 20, 5, seed=3)
    At = torch.sparse_csr_tensor(torch.tensor(A.indptr), torch.tensor(A.indices), torch.tensor(A.data),
                                 size=A.shape)
    x_true = 2e7 + 1e5 * np.random.default_rng(0).standard_normal(A.shape[0])
    b = torch.tensor(A @ x_true)
    x = amg_solve(At, b, tol=1e-10)
    assert x.device == b.device and x.dtype == b.dtype
    xs = spl.spsolve(A.tocsc(), b.numpy())
    assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-7
    assert amg_solve.last_stats["converged"]
    x2 = amg_solve(At, b * 1.5, tol=1e-10)  # same pattern: cached hierarchy, numeric update
    assert np.linalg.norm(x2.numpy() - 1.5 * xs) / np.linalg.norm(1.5 * xs) < 1e-7
    pre = amg_preconditioner(At)
    r = torch.tensor(np.random.default_rng(1).standard_normal(A.shape[0]), device=pre["device"])
    z = pre["apply"](r)
    assert torch.isfinite(z).all() and z.shape == r.shape


@pytest.mark.parametrize("hierarchy", ["default", "reference"])
def test_classical_amg_levels_metadata(hierarchy):
    """ClassicalAMG.levels carries the reference's keys (classical_amg.py:891-904,
    :925-933), each checked against its definition recomputed on the host from
    the exported A, P and C/F markers."""
    from matrices import impes_like
    from solver.classical_amg import ClassicalAMG

    A, *_ = impes_like(16, 16, 6, seed=5)
    At = torch.sparse_csr_tensor(torch.tensor(A.indptr), torch.tensor(A.indices), torch.tensor(A.data),
                                 size=A.shape)
    amg = ClassicalAMG(At, coarsest_size=50, hierarchy=hierarchy)
    lv = amg.levels
    assert len(lv) >= 3 and lv[-1]["is_coarsest"] and not any(d["is_coarsest"] for d in lv[:-1])
    coords = np.arange(A.shape[0], dtype=np.float64)
    for d in lv:
        Al = d["A"].to_dense().cpu().numpy()
        assert d["n"] == Al.shape[0] and d["basis"] is None
        assert np.array_equal(d["diag"].cpu().numpy(), np.abs(np.diag(Al)))
        if d["is_coarsest"]:
            assert d["P"] is None
            continue
        P = d["P"].to_dense().cpu().numpy()
        cf = d["cf_marker"].cpu().numpy()
        c, f = np.flatnonzero(cf == 0), np.flatnonzero(cf != 0)
        assert np.array_equal(d["coarse_nodes"].cpu().numpy(), c) and np.array_equal(d["F_nodes"].cpu().numpy(), f)
        f2c = d["fine2coarse"].cpu().numpy()
        assert np.array_equal(f2c[c], np.arange(c.size)) and np.all(f2c[f] == -1)
        assert np.array_equal(P[c], np.eye(P.shape[1])[f2c[c]])  # C rows inject
        assert np.array_equal(d["node_coords"].cpu().numpy(), coords)
        m = d["interp_metrics"]
        rs = P.sum(axis=1)
        assert np.isclose(m["row_err_max"], np.max(np.abs(rs - 1)), rtol=1e-12, atol=1e-15)
        assert np.isclose(m["const_rel"], np.linalg.norm(rs - 1) / np.sqrt(rs.size), rtol=1e-12, atol=1e-15)
        lin = P @ coords[c] - coords
        assert np.isclose(m["grad_rel"], np.linalg.norm(lin) / np.linalg.norm(coords), rtol=1e-12)
        w = d["P"].values().cpu().numpy()
        assert m["neg_count"] == int(np.sum(w < -1e-12)) and m["pos_count"] == int(np.sum(w > 1e-12))
        assert m["min_w"] == w.min() and m["max_w"] == w.max()
        coords = coords[c]


def test_amg_solve_reference_hierarchy():
    """amg_solve with the reference's hierarchy choices converges to the same
    answer as with the default one (they change iteration counts only), each
    cached under its own key."""
    import scipy.sparse.linalg as spl

    from matrices import impes_like
    from solver import pressure_amg as pa

    A, *_ = impes_like(20, 20, 5, seed=4)
    At = torch.sparse_csr_tensor(torch.tensor(A.indptr), torch.tensor(A.indices), torch.tensor(A.data),
                                 size=A.shape)
    b = torch.tensor(A @ (2e7 + 1e5 * np.random.default_rng(2).standard_normal(A.shape[0])))
    xs = spl.spsolve(A.tocsc(), b.numpy())
    its = {}
    for mode in ("default", "reference"):
        x = pa.amg_solve(At, b, tol=1e-10, hierarchy=mode, maxiter=600)
        assert pa.amg_solve.last_stats["converged"], mode
        its[mode] = pa.amg_solve.last_stats["iterations"]
        assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-7, mode
    assert sum(1 for k in pa._AMG_CACHE if k[0] == tuple(A.shape) and k[-1] in ("default", "reference")) >= 2
    with pytest.raises(ValueError):
        pa.amg_solve(At, b, hierarchy="other")


def test_impes_3phase_smoke():
    """tests/test_impes_3phase.py:58-73: 5 daily steps, no NaN, bounds, sum of saturations."""
    cfg = gas3_config()
    cfg["simulation"].update({"solver_type": "impes", "total_time_days": 5, "time_step_days": 1.0})
    _, _, fl, sim = mk(cfg)
    for _ in range(5):
        assert sim.run_step(dt=86400.0)
    assert not torch.isnan(fl.pressure).any() and not torch.isnan(fl.s_w).any() and not torch.isnan(fl.s_g).any()
    assert torch.all(fl.s_w >= fl.sw_cr) and torch.all(fl.s_w <= 1.0 - fl.so_r)
    assert torch.all(fl.s_g >= fl.sg_cr) and torch.all(fl.s_g <= 1.0 - fl.so_r - fl.s_w + 1e-8)
    s = fl.s_w + fl.s_g + fl.s_o
    assert torch.allclose(s, torch.ones_like(s), atol=1e-6)


def test_impes_3phase_mass_balance():
    """tests/test_impes_3phase.py:76-94: 10 daily steps, dM vs net inflow."""
    cfg = gas3_config()
    cfg["simulation"].update({"solver_type": "impes", "total_time_days": 10, "time_step_days": 1.0})
    res, _, fl, sim = mk(cfg)
    m0 = masses(res, fl)
    for _ in range(10):
        assert sim.run_step(dt=86400.0)
    m1 = masses(res, fl)
    mb = sim.mass_balance
    net = sum(mb[ph]["in"] - mb[ph]["out"] for ph in ("water", "gas", "oil"))
    atol = max(1e-6, 1e-3 * max(abs(m0), 1.0))
    assert abs((m1 - m0) - net) <= 5 * atol


def test_stone2_reduces_to_two_phase_when_sg_zero():
    """tests/test_relperm_stone2.py:11-38."""
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir

    res = Reservoir({"dimensions": [4, 4, 1], "grid_size": [10.0, 10.0, 10.0], "porosity": 0.2,
                     "permeability": 100.0, "k_vertical_fraction": 1.0}, torch.device("cuda"))
    fl = Fluid({"pressure": 20.0, "s_w": 0.3, "s_g": 0.0,
                "relative_permeability": {"model": "stone2", "nw": 2.0, "no": 2.0, "ng": 2.0, "ko_end_w": 1.0,
                                          "ko_end_g": 1.0}}, res, torch.device("cuda"))
    s_w = torch.full(res.dimensions, 0.6, dtype=torch.float64, device="cuda")
    fl.s_w = s_w
    fl.s_g = torch.zeros_like(s_w)
    kro_stone = fl.calc_oil_kr(s_w, fl.s_g)
    s_norm = (s_w - fl.sw_cr) / (1 - fl.sw_cr - fl.so_r)
    kro_corey = torch.clamp(1 - s_norm, 0.0, 1.0) ** fl.no
    assert torch.allclose(kro_stone, kro_corey, atol=1e-6)


def test_stone2_bounds():
    """tests/test_relperm_stone2.py:41-58."""
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir

    res = Reservoir({"dimensions": [2, 2, 1], "grid_size": [10, 10, 10], "porosity": 0.2, "permeability": 100.0,
                     "k_vertical_fraction": 1.0}, torch.device("cuda"))
    fl = Fluid({"pressure": 20.0, "s_w": 0.2, "s_g": 0.1, "relative_permeability": {"model": "stone2"}}, res,
               torch.device("cuda"))
    fl.s_w = torch.full(res.dimensions, 1.0 - fl.so_r - 1e-6, dtype=torch.float64, device="cuda")
    fl.s_g = torch.zeros_like(fl.s_w)
    assert (fl.calc_oil_kr(fl.s_w, fl.s_g) <= 1e-3).all()


def _box(p, sw, sg, wells, poro=0.2, perm=100.0):
    return {"simulation": {"solver_type": "impes", "total_time_days": 1, "time_step_days": 1},
            "reservoir": {"dimensions": [5, 5, 1], "grid_size": [10, 10, 10], "porosity": poro, "permeability": perm,
                          "k_vertical_fraction": 1.0},
            "fluid": {"pressure": p, "s_w": sw, "s_g": sg, "pvt_path": PVT}, "wells": wells}


def test_pcog_smoke_runs():
    """tests/test_pcog_smoke.py:13-44."""
    cfg = _box(20.0, 0.2, 0.05, [
        {"name": "INJ-G", "type": "injector", "i": 1, "j": 2, "k": 0, "radius": 0.1, "control_type": "rate",
         "control_value": 10.0, "injected_phase": "gas"},
        {"name": "PROD", "type": "producer", "i": 3, "j": 2, "k": 0, "radius": 0.1, "control_type": "bhp",
         "control_value": 18.0}])
    cfg["fluid"].pop("pvt_path")
    cfg["fluid"].update({"relative_permeability": {"model": "stone2"},
                         "capillary_pressure_og": {"pc_scale": 1e5, "pc_exponent": 1.5}})
    _, _, fl, sim = mk(cfg)
    assert sim.run_step(86400.0)
    assert torch.all(fl.s_w >= fl.sw_cr) and torch.all(fl.s_g >= fl.sg_cr)
    assert torch.all(fl.s_w + fl.s_g <= 1.0 - fl.so_r + 1e-6)


def test_wopr_limit_throttles_surface_oil_producer():
    """tests/test_phase_limits_and_autoswitch.py:11-33."""
    cfg = _box(20.0, 0.3, 0.02, [
        {"name": "PROD", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1, "control_type": "rate",
         "control_value": 50000.0, "rate_type": "surface", "surface_phase": "oil", "limits": {"wopr": 100.0}}])
    _, wells, _, sim = mk(cfg)
    assert sim.run_step(86400.0)
    assert wells.get_wells()[0].last_surface_rates.get("oil", 0.0) <= 100.0 * 1.10


def test_bhp_min_autoswitch_reduces_outflow():
    """tests/test_phase_limits_and_autoswitch.py:36-71."""
    q = []
    for extra in ({}, {"bhp_min": 9.5}):
        w = dict({"name": "PROD", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1, "control_type": "rate",
                  "control_value": 20000.0}, **extra)
        _, wells, _, sim = mk(_box(10.0, 0.2, 0.0, [w]))
        sim.run_step(86400.0)
        q.append(abs(wells.get_wells()[0].last_q_total or 0.0))
    assert q[1] <= q[0] + 1e-9


def test_component_balance_with_rv_surface_producer_and_gas_injector():
    """tests/test_component_balance_rv.py:11-40."""
    cfg = _box(20.0, 0.25, 0.05, [
        {"name": "INJ-G", "type": "injector", "i": 1, "j": 1, "k": 0, "radius": 0.1, "control_type": "rate",
         "control_value": 500.0, "rate_type": "surface", "injected_phase": "gas"},
        {"name": "PROD", "type": "producer", "i": 3, "j": 3, "k": 0, "radius": 0.1, "control_type": "rate",
         "control_value": 300.0, "rate_type": "surface", "surface_phase": "liquid"}], poro=0.25, perm=200.0)
    _, _, _, sim = mk(cfg)
    assert sim.run_step(86400.0)
    cb = sim.component_balance
    for ph in ("oil", "gas"):
        r = (cb[ph]["in"] - cb[ph]["out"]) - cb[ph]["accum"]
        assert abs(r) < 1e-2 * (1.0 + abs(cb[ph]["accum"]))


def test_surface_water_injection_mass_increase():
    """tests/test_surface_rate_conversion.py:13-66."""
    cfg = _box(20.0, 0.2, 0.0, [
        {"name": "INJ-W", "type": "injector", "i": 2, "j": 2, "k": 0, "radius": 0.1, "control_type": "rate",
         "control_value": 100.0, "injected_phase": "water", "rate_type": "surface"}])
    res, _, fl, sim = mk(cfg)
    V = res.cell_volume
    mw0 = torch.sum(res.porosity * fl.s_w * fl.rho_w * V).item()
    s_w0, rho_w0, phi = fl.s_w.clone(), fl.rho_w.clone(), res.porosity.clone()
    sim.run_step(86400.0)
    bw = float(fl._eval_pvt(fl.pressure, "Bw")[2, 2, 0])
    expected = 100.0 / 86400.0 * bw * float(fl.rho_w[2, 2, 0]) * 86400.0
    cap = float(torch.sum(phi * torch.clamp(1.0 - float(fl.so_r) - s_w0, min=0.0) * rho_w0) * V)
    eff = min(expected, cap)
    delta = torch.sum(res.porosity * fl.s_w * fl.rho_w * V).item() - mw0
    assert delta > 0
    assert abs(delta - eff) / max(eff, 1e-6) < 0.35


def test_pvt_rv_loading_and_fluid_rho_g(tmp_path):
    """tests/test_pvt_rv.py:12-54."""
    from simulator.fluid import Fluid
    from simulator.pvt import PVTTable
    from simulator.reservoir import Reservoir

    data = {"pressure_MPa": [10, 20, 40, 60], "Bo": [1.2, 1.1, 1.05, 1.03], "Bw": [1.02, 1.01, 1.00, 0.99],
            "Bg": [0.01, 0.005, 0.003, 0.002], "mu_o_cP": [2.0, 1.5, 1.3, 1.2], "mu_w_cP": [0.5, 0.52, 0.53, 0.54],
            "mu_g_cP": [0.02, 0.025, 0.03, 0.035], "Rs_m3m3": [80, 120, 150, 160],
            "Rv_m3m3": [0.0, 0.01, 0.015, 0.02], "units": {"pressure": "MPa"}}
    path = os.path.join(tmp_path, "pvt_rv.json")
    with open(path, "w", encoding="utf-8") as f:
        json.dump(data, f)
    out = PVTTable(path).eval(np.array([15.0, 55.0]))
    assert "Rv" in out and (out["Rv"] >= 0).all()
    res = Reservoir({"dimensions": [2, 2, 1], "grid_size": [10, 10, 10], "porosity": 0.2, "permeability": 100.0,
                     "k_vertical_fraction": 1.0}, torch.device("cuda"))
    fl = Fluid({"pressure": 40.0, "s_w": 0.2, "s_g": 0.1, "pvt_path": path}, res, torch.device("cuda"))
    assert torch.isfinite(fl.rho_g).all() and (fl.rho_g > 0).all()


def test_run_writes_the_reference_artifacts(tmp_path):
    """simulation.py:2120-2248: run() -> <name>.txt summary + mass balance, <name>.npz fields."""
    import re

    cfg = gas3_config()
    cfg["simulation"].update({"solver_type": "impes", "total_time_days": 4, "time_step_days": 1.0,
                              "save_interval": 2})
    _, _, fl, sim = mk(cfg)
    out = sim.run("case", results_root=str(tmp_path))
    txt = open(os.path.join(out, "case.txt")).read()
    assert txt.startswith("=== SUMMARY (final) ===\nPressure MPa: min=")
    assert re.search(r"^water: in=\S+, out=\S+, net=\S+$", txt, re.M)
    assert re.search(r"^total net=\S+$", txt, re.M)
    z = np.load(os.path.join(out, "case.npz"))
    assert set(z.files) == {"pressure", "sw", "sg", "so"}
    assert np.array_equal(z["pressure"], fl.pressure.double().cpu().numpy())
    # (i + 1) % save_interval == 0 and save_interval < num_steps: steps 2 and 4
    assert sorted(os.listdir(os.path.join(out, "intermediate"))) == ["case_step_2.npz", "case_step_4.npz"]
    mb = sim.mass_balance
    net = sum(mb[ph]["in"] - mb[ph]["out"] for ph in ("water", "gas", "oil"))
    assert f"total net={net:.6e}" in txt


def test_run_artifacts_match_oracle_run(tmp_path):
    """run() (simulation.py:2197-2248) on the gas3 case, 4 daily steps with the
    benchmark's solver settings, against an oracle run of the same case (exact
    pressure solves, its own state propagated): the final and intermediate npz fields (pressure 1e-7 relative,
    saturations 1e-6 absolute) and the results.txt figures (summary lines
    identical, mass-balance in/out/net and total 1e-6 relative against the
    oracle's cumulative mass ledger)."""
    import re

    from oracle import impes as oi

    from simulator.cases import BENCH_AMG

    cfg = gas3_config()
    cfg["simulation"].update({"solver_type": "impes", "total_time_days": 4, "time_step_days": 1.0,
                              "save_interval": 2, "pressure_solver": "amg", "amg": dict(BENCH_AMG)})
    _, _, fl, sim = mk(cfg)
    out = sim.run("case", results_root=str(tmp_path))
    orc = oi.Simulation(oi.Case(cfg))
    states = []
    for _ in range(4):
        orc.step(86400.0)
        st = orc.state
        states.append({"pressure": st["p"], "sw": st["sw"], "sg": st["sg"], "so": 1.0 - st["sw"] - st["sg"]})

    def check_fields(z, ref):
        assert np.max(np.abs(z["pressure"] - ref["pressure"])) <= 1e-7 * np.max(np.abs(ref["pressure"]))
        for k in ("sw", "sg", "so"):
            assert np.max(np.abs(z[k] - ref[k])) <= 1e-6, k

    check_fields(np.load(os.path.join(out, "case.npz")), states[-1])
    for step in (2, 4):
        check_fields(np.load(os.path.join(out, "intermediate", f"case_step_{step}.npz")), states[step - 1])
    txt = open(os.path.join(out, "case.txt")).read()
    p, sw, sg, so = (states[-1][k] for k in ("pressure", "sw", "sg", "so"))
    expect = [f"Pressure MPa: min={p.min()/1e6:.2f}, max={p.max()/1e6:.2f}, mean={p.mean()/1e6:.2f}",
              f"Sw: min={sw.min():.4f}, max={sw.max():.4f}, mean={sw.mean():.4f}",
              f"Sg: min={sg.min():.4f}, max={sg.max():.4f}, mean={sg.mean():.4f}",
              f"So: min={so.min():.4f}, max={so.max():.4f}, mean={so.mean():.4f}"]
    for line in expect:
        assert line in txt, line
    mb = orc.mass_balance
    total = 0.0
    for ph in ("water", "gas", "oil"):
        m = re.search(rf"^{ph}:\s+in=(\S+), out=(\S+), net=(\S+)$", txt, re.M)
        vin, vout, net = (float(v) for v in m.groups())
        tol = 1e-6 * (abs(mb[ph]["in"]) + abs(mb[ph]["out"])) + 1e-9
        assert abs(vin - mb[ph]["in"]) <= tol and abs(vout - mb[ph]["out"]) <= tol, ph
        assert abs(net - (mb[ph]["in"] - mb[ph]["out"])) <= tol, ph
        total += mb[ph]["in"] - mb[ph]["out"]
    m = re.search(r"^total net=(\S+)$", txt, re.M)
    scale = sum(abs(mb[ph]["in"]) + abs(mb[ph]["out"]) for ph in mb)
    assert abs(float(m.group(1)) - total) <= 1e-6 * scale


def test_update_matrix_rejects_a_size_change():
    """ClassicalAMG.update_matrix raises ValueError when the new matrix has
    another size (reference classical_amg.py:1366-1367) and keeps working for
    a same-size update."""
    from matrices import impes_like
    from solver.classical_amg import ClassicalAMG

    def csr(A, dev):
        return torch.sparse_csr_tensor(torch.tensor(A.indptr), torch.tensor(A.indices), torch.tensor(A.data),
                                       size=A.shape).to(dev)

    A1, *_ = impes_like(10, 8, 4, seed=2)
    A2, *_ = impes_like(10, 8, 5, seed=2)
    amg = ClassicalAMG(csr(A1, "cuda"))
    with pytest.raises(ValueError):
        amg.update_matrix(csr(A2, "cuda"))
    A3, *_ = impes_like(10, 8, 4, seed=2, sigma=1.3)
    amg.update_matrix(csr(A3, "cuda"))
    b = torch.ones(A1.shape[0], dtype=torch.float64, device="cuda")
    x = amg.solve(b, tol=1e-6, max_iter=50)
    assert torch.isfinite(x).all()
