"""Writes the golden vectors of tests/golden/*.npz from the reference's own code.

Run in the build container only (the reference is not present on GPU hosts):
    python tests/golden/make_golden.py [--reference /root/reference]

For each case the reference Simulator runs a few IMPES steps with
``pressure_solver = "amg"`` while ``solver.pressure_amg.amg_solve`` is replaced
by an exact FP64 sparse solve that records the assembled system. Per step the
fixture stores the state before the step, the reference's pressure matrix (COO)
and right-hand side, the pressure it used, the state after its transport step,
its substep count and its ledgers and per-well diagnostics. A second fixture
holds reference fluid-property evaluations (relative permeabilities, capillary
pressure, PVT lookups, compressibility, densities) on sample inputs.
The reference computes in float32; the tests compare at float32 tolerance.
"""
from __future__ import annotations

import argparse
import contextlib
import io
import json
import os
import re
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))


def cases(ref: str) -> dict:
    pvt = os.path.join(ref, "configs", "pvt", "pvt_synthetic.json")
    with open(os.path.join(ref, "configs", "gas_injection_3phase.json")) as f:
        gas3 = json.load(f)
    gas3["simulation"].update({"time_step_days": 1.0})
    rv_pvt = {
        "pressure_MPa": [10, 20, 40, 60], "Bo": [1.2, 1.1, 1.05, 1.03], "Bw": [1.02, 1.01, 1.00, 0.99],
        "Bg": [0.01, 0.005, 0.003, 0.002], "mu_o_cP": [2.0, 1.5, 1.3, 1.2], "mu_w_cP": [0.5, 0.52, 0.53, 0.54],
        "mu_g_cP": [0.02, 0.025, 0.03, 0.035], "Rs_m3m3": [80, 120, 150, 160], "Rv_m3m3": [0.0, 0.01, 0.015, 0.02],
        "units": {"pressure": "MPa"},
    }
    res5 = {"dimensions": [5, 5, 1], "grid_size": [10, 10, 10], "porosity": 0.2, "permeability": 100.0,
            "k_vertical_fraction": 1.0}
    sim1 = {"solver_type": "impes"}
    out = {
        # tests/test_impes_3phase.py, test_amg_pressure_solver.py config (gravity, 2 BHP producers)
        "gas3": dict(cfg=gas3, steps=3, dt=1.0),
        # tests/test_surface_rate_conversion.py
        "surf_water_inj": dict(cfg={"simulation": sim1, "reservoir": res5,
                                    "fluid": {"pressure": 20.0, "s_w": 0.2, "s_g": 0.0, "pvt_path": pvt},
                                    "wells": [{"name": "INJ-W", "type": "injector", "i": 2, "j": 2, "k": 0,
                                               "radius": 0.1, "control_type": "rate", "control_value": 100.0,
                                               "injected_phase": "water", "rate_type": "surface"}]},
                               steps=2, dt=1.0),
        # tests/test_component_balance_rv.py
        "comp_rv": dict(cfg={"simulation": sim1,
                             "reservoir": dict(res5, porosity=0.25, permeability=200.0),
                             "fluid": {"pressure": 20.0, "s_w": 0.25, "s_g": 0.05, "pvt_path": pvt},
                             "wells": [{"name": "INJ-G", "type": "injector", "i": 1, "j": 1, "k": 0, "radius": 0.1,
                                        "control_type": "rate", "control_value": 500.0, "rate_type": "surface",
                                        "injected_phase": "gas"},
                                       {"name": "PROD", "type": "producer", "i": 3, "j": 3, "k": 0, "radius": 0.1,
                                        "control_type": "rate", "control_value": 300.0, "rate_type": "surface",
                                        "surface_phase": "liquid"}]},
                        steps=2, dt=1.0),
        # tests/test_phase_limits_and_autoswitch.py (wopr limit)
        "wopr_limit": dict(cfg={"simulation": sim1, "reservoir": res5,
                                "fluid": {"pressure": 20.0, "s_w": 0.3, "s_g": 0.02, "pvt_path": pvt},
                                "wells": [{"name": "PROD", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1,
                                           "control_type": "rate", "control_value": 50000.0, "rate_type": "surface",
                                           "surface_phase": "oil", "limits": {"wopr": 100.0}}]},
                           steps=1, dt=1.0),
        # tests/test_phase_limits_and_autoswitch.py (bhp_min autoswitch)
        "bhp_min": dict(cfg={"simulation": sim1, "reservoir": res5,
                             "fluid": {"pressure": 10.0, "s_w": 0.2, "s_g": 0.0, "pvt_path": pvt},
                             "wells": [{"name": "PROD2", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1,
                                        "control_type": "rate", "control_value": 20000.0, "bhp_min": 9.5}]},
                        steps=1, dt=1.0),
        # tests/test_pcog_smoke.py (Stone II + oil-gas capillarity, BHP producer)
        "pcog": dict(cfg={"simulation": sim1, "reservoir": res5,
                          "fluid": {"pressure": 20.0, "s_w": 0.2, "s_g": 0.05,
                                    "relative_permeability": {"model": "stone2"},
                                    "capillary_pressure_og": {"pc_scale": 1e5, "pc_exponent": 1.5}},
                          "wells": [{"name": "INJ-G", "type": "injector", "i": 1, "j": 2, "k": 0, "radius": 0.1,
                                     "control_type": "rate", "control_value": 10.0, "injected_phase": "gas"},
                                    {"name": "PROD", "type": "producer", "i": 3, "j": 2, "k": 0, "radius": 0.1,
                                     "control_type": "bhp", "control_value": 18.0}]},
                     steps=2, dt=1.0),
        # tests/test_surface_oil_producer_with_rv.py (PVT with Rv)
        "oil_prod_rv": dict(cfg={"simulation": sim1, "reservoir": res5,
                                 "fluid": {"pressure": 20.0, "s_w": 0.3, "s_g": 0.05, "pvt_path": "@rv"},
                                 "wells": [{"name": "PROD", "type": "producer", "i": 2, "j": 2, "k": 0,
                                            "radius": 0.1, "control_type": "rate", "control_value": 100.0,
                                            "rate_type": "surface", "surface_phase": "oil"}]},
                            steps=2, dt=1.0, pvt_inline=rv_pvt),
        # tests/test_surface_gas_producer_conversion.py
        "gas_prod": dict(cfg={"simulation": sim1, "reservoir": res5,
                              "fluid": {"pressure": 20.0, "s_w": 0.3, "s_g": 0.05, "pvt_path": pvt},
                              "wells": [{"name": "PROD", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1,
                                         "control_type": "rate", "control_value": 200.0, "rate_type": "surface",
                                         "surface_phase": "gas"}]},
                         steps=2, dt=1.0),
        # 3D, gravity, o-w capillarity, Corey, no PVT, rate wells (configs/test_config_3d.json + fast_test Pc)
        "grav_pc_3d": dict(cfg={"simulation": sim1,
                                "reservoir": {"dimensions": [6, 5, 4], "grid_size": [25.0, 25.0, 15.0],
                                              "porosity": 0.25, "permeability": 50, "k_vertical_fraction": 0.3},
                                "fluid": {"pressure": 15.0, "s_w": 0.2, "mu_oil": 2.0, "mu_water": 1.0,
                                          "rho_oil": 800.0, "rho_water": 1000.0, "c_oil": 3e-5, "c_water": 3e-5,
                                          "c_rock": 3e-5,
                                          "relative_permeability": {"sw_cr": 0.1, "so_r": 0.2, "nw": 2.0, "no": 2.0},
                                          "capillary_pressure": {"pc_scale": 1e4, "pc_exponent": 1.5}},
                                "wells": [{"name": "P1", "type": "producer", "coordinates": [1, 1, 1], "radius": 0.1,
                                           "control": {"type": "rate", "value": 30.0}},
                                          {"name": "I1", "type": "injector", "coordinates": [4, 3, 2], "radius": 0.1,
                                           "control": {"type": "rate", "value": 30.0}}]},
                           steps=3, dt=1.0),
        # two-phase oil-water, 2D, constant compressibility (the shape of the 60x60 benchmark case)
        "ow_2d": dict(cfg={"simulation": sim1,
                           "reservoir": {"dimensions": [12, 10, 1], "grid_size": [10.0, 10.0, 5.0], "porosity": 0.2,
                                         "permeability": 100.0},
                           "fluid": {"pressure": 20.0, "s_w": 0.2, "mu_oil": 5.0, "mu_water": 1.0,
                                     "c_oil": 7.5e-4, "c_water": 7.5e-4, "c_gas": 7.5e-4, "c_rock": 7.5e-4,
                                     "relative_permeability": {"nw": 2, "no": 2, "sw_cr": 0.2, "so_r": 0.2}},
                           "wells": [{"name": "INJ", "type": "injector", "i": 0, "j": 0, "k": 0, "radius": 0.1,
                                      "control_type": "rate", "control_value": 100.0, "injected_phase": "water"},
                                     {"name": "PROD", "type": "producer", "i": 11, "j": 9, "k": 0, "radius": 0.1,
                                      "control_type": "rate", "control_value": 80.0}]},
                      steps=4, dt=1.0),
    }
    return out


def run_case(name, spec, tmpdir):
    import torch
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir
    from simulator.simulation import Simulator
    from simulator.well import WellManager
    import solver.pressure_amg as pa

    cfg = json.loads(json.dumps(spec["cfg"]))
    pvt_data = None
    if cfg["fluid"].get("pvt_path") == "@rv":
        path = os.path.join(tmpdir, f"{name}_pvt.json")
        with open(path, "w") as f:
            json.dump(spec["pvt_inline"], f)
        cfg["fluid"]["pvt_path"] = path
        pvt_data = spec["pvt_inline"]
    elif cfg["fluid"].get("pvt_path"):
        with open(cfg["fluid"]["pvt_path"]) as f:
            pvt_data = json.load(f)
    cfg["simulation"]["solver_type"] = "impes"
    cfg["simulation"]["pressure_solver"] = "amg"
    rec = {}

    def exact(A, b, **kw):
        A = A.coalesce()
        idx = A.indices().numpy()
        As = sp.csr_matrix((A.values().double().numpy(), (idx[0], idx[1])), shape=tuple(A.shape))
        bb = b.double().numpy()
        x = spl.spsolve(As.tocsc(), bb)
        rec["A_row"], rec["A_col"], rec["A_val"] = idx[0].astype(np.int32), idx[1].astype(np.int32), A.values().numpy()
        rec["b"] = b.numpy().copy()
        xt = torch.tensor(x, dtype=b.dtype)
        rec["p_new"] = xt.numpy().copy()
        return xt

    pa.amg_solve = exact
    dev = torch.device("cpu")
    with contextlib.redirect_stdout(io.StringIO()):
        res = Reservoir(cfg["reservoir"], dev)
        wm = WellManager(cfg["wells"], res)
        fl = Fluid(cfg["fluid"], res, dev)
        sim = Simulator(res, fl, wm, cfg["simulation"], dev)
    steps = []
    for s in range(spec["steps"]):
        before = {"p": fl.pressure.numpy().copy(), "sw": fl.s_w.numpy().copy(), "sg": fl.s_g.numpy().copy()}
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            ok = sim.run_step(spec["dt"] * 86400.0)
        assert ok
        m = re.findall(r"substeps N=(\d+)", buf.getvalue())
        steps.append({
            "before": before, "A_row": rec["A_row"], "A_col": rec["A_col"], "A_val": rec["A_val"], "b": rec["b"],
            "p_new": rec["p_new"], "sw": fl.s_w.numpy().copy(), "sg": fl.s_g.numpy().copy(),
            "N": int(m[-1]) if m else -1,
            "mass": json.loads(json.dumps(sim.mass_balance)),
            "comp": json.loads(json.dumps(sim.component_balance, default=float)),
            "wells": [{"name": w.name, "last_q_total": None if w.last_q_total is None else float(w.last_q_total),
                       "last_mode": w.last_mode,
                       "last_surface_rates": {k: float(v) for k, v in (w.last_surface_rates or {}).items()}}
                      for w in wm.get_wells()],
        })
    return cfg, pvt_data, steps


def fluid_goldens(ref):
    import torch
    from simulator.fluid import Fluid
    from simulator.reservoir import Reservoir

    pvt = os.path.join(ref, "configs", "pvt", "pvt_synthetic.json")
    out = {}
    dev = torch.device("cpu")
    sw = np.linspace(0.0, 1.0, 41, dtype=np.float32)
    sg = np.linspace(0.0, 0.8, 41, dtype=np.float32)
    SW, SG = np.meshgrid(sw, sg, indexing="ij")
    p = np.linspace(1e6, 70e6, 41 * 41, dtype=np.float32).reshape(41, 41)
    variants = {
        "corey": {"relative_permeability": {"nw": 2.0, "no": 2.5, "ng": 2.2, "sw_cr": 0.15, "so_r": 0.2, "sg_cr": 0.02},
                  "capillary_pressure": {"pc_scale": 1e4, "pc_exponent": 1.5}},
        "stone2": {"relative_permeability": {"model": "stone2", "nw": 2.0, "no": 2.0, "ng": 2.0, "now": 2.5,
                                             "nog": 1.8, "ko_end_w": 0.9, "ko_end_g": 0.8, "sw_cr": 0.2,
                                             "so_r": 0.15, "sg_cr": 0.05},
                   "capillary_pressure_og": {"pc_scale": 1e5, "pc_exponent": 1.5}},
        "pvt": {"pvt_path": pvt},
        "nopvt": {"c_oil": 1e-5, "c_water": 5e-6, "c_gas": 1e-3, "rho_gas": 1.2},
    }
    for vname, extra in variants.items():
        res_cfg = {"dimensions": [41, 41, 1], "grid_size": [10, 10, 10], "porosity": 0.2, "permeability": 100.0}
        with contextlib.redirect_stdout(io.StringIO()):
            res = Reservoir(res_cfg, dev)
            fl = Fluid(dict({"pressure": 20.0, "s_w": 0.2, "s_g": 0.0}, **extra), res, dev)
        tsw = torch.tensor(SW[..., None])
        tsg = torch.tensor(SG[..., None])
        fl.s_g = tsg
        kro, krw, krg = fl.get_rel_perms(tsw)
        out[f"{vname}_kro"], out[f"{vname}_krw"], out[f"{vname}_krg"] = kro.numpy(), krw.numpy(), krg.numpy()
        out[f"{vname}_pc"] = fl.get_capillary_pressure(tsw).numpy()
        out[f"{vname}_pcog"] = fl.get_capillary_pressure_og(tsg).numpy()
        tp = torch.tensor(p[..., None])
        fl.pressure = tp
        out[f"{vname}_ct"] = fl.calc_total_compressibility(tp, tsw, tsg).numpy()
        out[f"{vname}_rho_w"], out[f"{vname}_rho_o"], out[f"{vname}_rho_g"] = (
            torch.as_tensor(fl.rho_w).numpy(), torch.as_tensor(fl.rho_o).numpy(), torch.as_tensor(fl.rho_g).numpy())
        if fl.pvt is not None:
            for key in ("Bo", "Bw", "Bg", "mu_o", "mu_w", "mu_g", "Rs", "Rv"):
                out[f"{vname}_pvt_{key}"] = fl._eval_pvt(tp, key).numpy()
    out["SW"], out["SG"], out["P"] = SW, SG, p
    cfgs = {k: v for k, v in variants.items()}
    return out, cfgs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    ref = os.path.abspath(a.reference)
    sys.path.insert(0, os.path.join(ref, "src"))
    import tempfile

    meta = {}
    with tempfile.TemporaryDirectory() as td:
        for name, spec in cases(ref).items():
            cfg, pvt_data, steps = run_case(name, spec, td)
            arrays = {}
            for s, st in enumerate(steps):
                for k in ("A_row", "A_col", "A_val", "b", "p_new", "sw", "sg"):
                    arrays[f"s{s}_{k}"] = st[k]
                for k, v in st["before"].items():
                    arrays[f"s{s}_before_{k}"] = v
            np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **arrays)
            cfg_out = json.loads(json.dumps(cfg))
            if cfg_out["fluid"].get("pvt_path"):
                cfg_out["fluid"]["pvt_path"] = None
            meta[name] = {"config": cfg_out, "pvt": pvt_data, "dt_days": spec["dt"],
                          "steps": [{k: st[k] for k in ("N", "mass", "comp", "wells")} for st in steps]}
    with open(os.path.join(HERE, "steps.json"), "w") as f:
        json.dump(meta, f, indent=1)
    fl, cfgs = fluid_goldens(ref)
    np.savez_compressed(os.path.join(HERE, "fluid.npz"), **fl)
    with open(os.path.join(HERE, "fluid.json"), "w") as f:
        json.dump(cfgs, f, indent=1)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
