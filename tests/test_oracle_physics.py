"""Closed forms, invariants and reference fluid-property goldens for the oracle. CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import impes as oi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FL = np.load(os.path.join(GOLD, "fluid.npz"))
FLCFG = json.load(open(os.path.join(GOLD, "fluid.json")))
PVT_SYNTH = json.load(open(os.path.join(GOLD, "pvt_synthetic.json")))


def fluid_case(variant):
    fc = dict({"pressure": 20.0, "s_w": 0.2, "s_g": 0.0}, **FLCFG[variant])
    pvt = None
    if "pvt_path" in fc:
        fc.pop("pvt_path")
        pvt = oi.PVT(PVT_SYNTH)
    cfg = {"reservoir": {"dimensions": [41, 41, 1], "grid_size": [10, 10, 10]}, "fluid": fc, "wells": []}
    return oi.Case(cfg, pvt=pvt)


@pytest.mark.parametrize("variant", ["corey", "stone2", "pvt", "nopvt"])
def test_fluid_properties_match_reference(variant):
    c = fluid_case(variant)
    SW, SG, P = (FL[k].astype(np.float64)[..., None] for k in ("SW", "SG", "P"))
    kro, krw, krg = c.relperms(SW, SG)
    for name, got in (("kro", kro), ("krw", krw), ("krg", krg)):
        ref = FL[f"{variant}_{name}"].astype(np.float64)
        assert np.allclose(got, ref, rtol=2e-6, atol=1e-7), (variant, name)
    # capillary pressure: scale*(1 - s_n + 1e-6)^-e is ill-conditioned near s_n = 1, where the
    # reference's float32 rounding of (1 - s_n) dominates; compare where 1 - s_n > 1e-2 and
    # check the float32 emulation everywhere
    for name, sat, scale, ex, den, off in (
            ("pc", SW, c.pc_scale, c.pc_exponent, 1 - c.sw_cr - c.so_r, c.sw_cr),
            ("pcog", SG, c.pcog_scale, c.pcog_exponent, 1 - c.sw_cr - c.so_r + 1e-12, 0.0)):
        ref = FL[f"{variant}_{name}"].astype(np.float64)
        got = c.pc_ow(SW) if name == "pc" else c.pc_og(SG)
        sn = np.clip((sat - off) / den, 0, 1)
        # float32 noise of (1 - s_n) amplified by the exponent
        rtol = 2e-6 + 2e-7 * ex / (1 - sn + 1e-6)
        assert np.all(np.abs(got - ref) <= rtol * np.abs(ref) + 1e-9), (variant, name)
        if scale:
            sat32 = sat.astype(np.float32)
            sn32 = np.clip(((sat32 - np.float32(off)) / np.float32(den)).astype(np.float32), 0, 1)
            base = ((np.float32(1.0) - sn32) + np.float32(1e-6)).astype(np.float32)
            emu = (np.float32(scale) * base.astype(np.float64) ** (-ex)).astype(np.float32)
            assert np.allclose(emu, ref, rtol=5e-6), (variant, name)
    rho = c.densities(P)
    for name, got in zip(("rho_w", "rho_o", "rho_g"), rho):
        ref = FL[f"{variant}_{name}"].astype(np.float64)
        assert np.allclose(got, ref, rtol=2e-6), (variant, name)
    if c.pvt is not None:
        for key in ("Bo", "Bw", "Bg", "mu_o", "mu_w", "mu_g", "Rs", "Rv"):
            ref = FL[f"{variant}_pvt_{key}"].astype(np.float64)
            assert np.allclose(c.pvt(key, P), ref, rtol=2e-6, atol=1e-9), key


def test_stone2_reduces_to_corey_kat():
    """tests/test_relperm_stone2.py:11-56 (reference KAT)."""
    cfg = {"reservoir": {"dimensions": [4, 4, 1]}, "fluid": {"relative_permeability": {
        "model": "stone2", "nw": 2.0, "no": 2.0, "ng": 2.0, "ko_end_w": 1.0, "ko_end_g": 1.0}}, "wells": []}
    c = oi.Case(cfg)
    sw = np.full((4, 4, 1), 0.6)
    kro, _, _ = c.relperms(sw, np.zeros_like(sw))
    s_norm = (sw - 0.2) / (1 - 0.2 - 0.2)
    assert np.allclose(kro, np.clip(1 - s_norm, 0, 1) ** 2, atol=1e-6)
    kro2, _, _ = c.relperms(np.full((4, 4, 1), 1 - 0.2 - 1e-6), np.zeros((4, 4, 1)))
    assert np.all(kro2 <= 1e-3)


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def test_compressibility_float32_emulation():
    """The reference's ct (fluid.py:190-222) in its own float32 arithmetic is the
    oracle's formula evaluated with float32 rounding at every torch op."""
    c = fluid_case("pvt")
    SW, SG, P = (FL[k][..., None] for k in ("SW", "SG", "P"))

    def B(key, x32):  # _eval_pvt(x32 * 1e6) with the float32 tensor ops of the reference
        pt = _f32(x32 * _f32(1e6))
        p_mpa = _f32(pt / _f32(1e6)).astype(np.float64)
        return _f32(np.interp(p_mpa, c.pvt.P, c.pvt.cols[key], left=c.pvt.cols[key][0], right=c.pvt.cols[key][-1]))

    pm = _f32(_f32(P) / _f32(1e6))
    pp, pn = _f32(pm + _f32(0.1)), _f32(pm - _f32(0.1))
    cc = {}
    for key in ("Bw", "Bo", "Bg"):
        b0, bp, bm = B(key, pm), B(key, pp), B(key, pn)
        dB = _f32(_f32(bp - bm) / _f32(0.2))
        cc[key] = _f32(-_f32(dB / _f32(b0 + _f32(1e-12))) / _f32(1e6))
    so = _f32(_f32(1.0) - SW - SG)
    ct32 = _f32(_f32(c.c_r) + SW * cc["Bw"]) + so * cc["Bo"] + SG * cc["Bg"]
    ref = FL["pvt_ct"]
    assert np.allclose(ct32, ref, rtol=1e-5, atol=1e-15)
    # the FP64 oracle differs from the float32 reference only by that rounding noise
    assert np.allclose(c.ct(P.astype(np.float64), SW.astype(np.float64), SG.astype(np.float64)), ref, rtol=3e-3,
                       atol=1e-13)


def test_peaceman_index_closed_form():
    w = oi.Well("W", "producer", 0, 0, 0, 0.1, "bhp", 15.0)
    ro = 0.28 * math.sqrt(2.0) / 2
    assert math.isclose(w.wi, 2 * math.pi * 100.0 / (math.log(ro / 0.1) + 1e-12), rel_tol=1e-15)
    assert math.isclose(w.wi, 919.8775630729912, rel_tol=1e-12)


def test_transmissibility_closed_form():
    cfg = {"reservoir": {"dimensions": [2, 1, 1], "grid_size": [10.0, 10.0, 5.0]}, "fluid": {}, "wells": []}
    c = oi.Case(cfg, fields={"perm_x": np.array([100.0, 300.0]), "perm_y": np.array([1.0, 1.0]),
                             "perm_z": np.array([1.0, 1.0])})
    k1, k2 = 100 * oi.MD_TO_M2, 300 * oi.MD_TO_M2
    assert math.isclose(c.Tx[0, 0, 0], 2 * k1 * k2 / (k1 + k2 + 1e-15) * 10 * 5 / 10, rel_tol=1e-14)
    # the 1e-15 m^2 epsilon biases a 100 mD face by -0.50 %
    k = 100 * oi.MD_TO_M2
    assert math.isclose((2 * k * k / (2 * k + 1e-15)) / k - 1, -0.00504, rel_tol=2e-2)


def _small_case(**over):
    cfg = {"simulation": {}, "reservoir": {"dimensions": [6, 5, 3], "grid_size": [10.0, 10.0, 4.0],
                                           "porosity": 0.2, "permeability": 80.0},
           "fluid": {"pressure": 20.0, "s_w": 0.3, "s_g": 0.05, "mu_oil": 2.0, "mu_water": 0.6,
                     "relative_permeability": {"model": "stone2", "sw_cr": 0.2, "so_r": 0.2, "sg_cr": 0.0}},
           "wells": []}
    for k, v in over.items():
        cfg[k].update(v) if isinstance(v, dict) else cfg.__setitem__(k, v)
    return oi.Case(cfg, pvt=oi.PVT(PVT_SYNTH))


def test_pressure_matrix_invariants():
    c = _small_case(wells=[{"name": "P", "type": "producer", "i": 2, "j": 2, "k": 1, "radius": 0.1,
                            "control_type": "bhp", "control_value": 18.0}])
    st = c.initial_state()
    st["p"] = st["p"] + 1e5 * np.random.default_rng(0).standard_normal(c.dims)
    ps = oi.pressure_system(c, st, 86400.0)
    A = oi.system_csr(c, ps).toarray()
    assert np.allclose(A, A.T, rtol=0, atol=0)
    off = A - np.diag(np.diag(A))
    assert np.all(off <= 0)
    # row sums = accumulation + well diagonal
    assert np.allclose(A.sum(axis=1), ps["acc"] + ps["bhp_diag"], rtol=1e-12, atol=1e-25)


def test_uniform_pressure_without_wells_keeps_saturation():
    c = _small_case()
    c.dz = 0.0  # no gravity: zero flux at uniform pressure
    st = c.initial_state()
    ps = oi.pressure_system(c, st, 86400.0)
    out = oi.saturation_step(c, st, st["p"], 86400.0, ps)
    assert np.array_equal(out["sw"], st["sw"]) and np.array_equal(out["sg"], st["sg"])


def test_bhp_row_balance_equals_peaceman_rate_when_well_conditioned():
    """The row-balance form of the BHP rate equals WI*lambda*(p - p_bhp) for an exact
    pressure solve; with a viscous fluid and permeable rock the direct form is
    itself accurate, so both must agree."""
    # p_cell - p_bhp = q/(WI*lambda) must be many ulp of p_cell: low absolute
    # pressure at the producer and a large injection rate
    c = _small_case(fluid={"pressure": 0.001, "s_g": 0.0},
                    wells=[{"name": "I", "type": "injector", "i": 0, "j": 0, "k": 0, "radius": 0.1,
                            "control_type": "rate", "control_value": 5e6},
                           {"name": "P", "type": "producer", "i": 5, "j": 4, "k": 2, "radius": 0.1,
                            "control_type": "bhp", "control_value": 0.001}])
    c.pvt = None
    st = c.initial_state()
    ps = oi.pressure_system(c, st, 86400.0)
    p = oi.solve_direct(oi.system_csr(c, ps), ps["b"])
    bal = oi.bhp_row_balance(c, ps, p, st["p"])
    w = c.wells[1]
    idx = c.flat(w)
    direct = ps["terms"][1] * (w.control_value * 1e6 - p[idx])
    assert abs(direct) > 1e-1
    # the direct form keeps ~ulp(p)/(p - p_bhp) ~ 1e-10 relative error here
    assert math.isclose(bal[idx], direct, rel_tol=1e-8)


def test_volume_ledger_closes_in_fp64():
    cfg = {"simulation": {}, "reservoir": {"dimensions": [8, 7, 3], "grid_size": [10.0, 10.0, 3.0],
                                           "porosity": 0.2, "permeability": 100.0},
           "fluid": {"pressure": 20.0, "s_w": 0.25, "s_g": 0.02,
                     "relative_permeability": {"model": "stone2"}},
           "wells": [{"name": "G", "type": "injector", "i": 0, "j": 0, "k": 0, "radius": 0.1,
                      "control_type": "rate", "control_value": 400.0, "rate_type": "surface", "injected_phase": "gas"},
                     {"name": "W", "type": "injector", "i": 7, "j": 0, "k": 2, "radius": 0.1,
                      "control_type": "rate", "control_value": 30.0},
                     {"name": "P", "type": "producer", "i": 4, "j": 6, "k": 1, "radius": 0.1,
                      "control_type": "bhp", "control_value": 18.0}]}
    sim = oi.Simulation(oi.Case(cfg, pvt=oi.PVT(PVT_SYNTH)))
    for _ in range(5):
        sim.step(86400.0)
    # the ledger closes even where the reference's explicit scheme has gone
    # unstable (substeps capped at 20 drive S_g above 1 here)
    err = sim.volume_balance_error()
    assert err["water"] < 1e-12 and err["gas"] < 1e-12, err


def test_saturation_bounds_mild_case():
    """tests/test_impes_3phase.py:60-73 invariants on a stable case."""
    cfg = {"simulation": {}, "reservoir": {"dimensions": [8, 7, 3], "grid_size": [10.0, 10.0, 3.0],
                                           "porosity": 0.2, "permeability": 100.0},
           "fluid": {"pressure": 20.0, "s_w": 0.25, "s_g": 0.02, "relative_permeability": {"model": "stone2"}},
           "wells": [{"name": "G", "type": "injector", "i": 0, "j": 0, "k": 0, "radius": 0.1,
                      "control_type": "rate", "control_value": 20.0, "injected_phase": "gas"},
                     {"name": "W", "type": "injector", "i": 7, "j": 0, "k": 2, "radius": 0.1,
                      "control_type": "rate", "control_value": 10.0},
                     {"name": "P", "type": "producer", "i": 4, "j": 6, "k": 1, "radius": 0.1,
                      "control_type": "rate", "control_value": 25.0}]}
    sim = oi.Simulation(oi.Case(cfg, pvt=oi.PVT(PVT_SYNTH)))
    for _ in range(5):
        sim.step(86400.0)
    s, c = sim.state, sim.case
    assert np.all(np.isfinite(s["p"]))
    assert np.all(s["sw"] >= c.sw_cr) and np.all(s["sw"] <= 1 - c.so_r)
    assert np.all(s["sg"] >= c.sg_cr) and np.all(s["sg"] <= 1 - c.so_r - s["sw"] + 1e-8)
    err = sim.volume_balance_error()
    assert err["water"] < 1e-12 and err["gas"] < 1e-12, err


@pytest.mark.parametrize("limit,phase", [("wopr", "oil"), ("wlpr", "water"), ("wgpr", "gas"), ("liqr", "liquid")])
def test_producer_rate_limits_cap_surface_rates(limit, phase):
    """simulation.py:1262-1320: a surface-rate producer whose control exceeds a
    limit is cut back so that the limited surface rate equals its cap while the
    other surface rates follow in proportion (same cell fractions); without the
    limit the same well produces its control. Surface rates are recomputed here
    from the oracle's reservoir rate and the cell's fractional flows and PVT
    values (the definitions of well.last_surface_rates). The cap is half the
    unlimited well's surface rate of that phase."""
    base = {"simulation": {"solver_type": "impes"},
            "reservoir": {"dimensions": [5, 5, 1], "grid_size": [10, 10, 10], "porosity": 0.25,
                          "permeability": 200.0, "k_vertical_fraction": 1.0},
            "fluid": {"pressure": 25.0, "s_w": 0.45, "s_g": 0.15, "pvt_table": PVT_SYNTH},
            "wells": [{"name": "P", "type": "producer", "i": 2, "j": 2, "k": 0, "radius": 0.1,
                       "control_type": "rate", "control_value": 8.0, "rate_type": "surface",
                       "surface_phase": "liquid"}]}
    out = {}
    cap = None
    for limited in (False, True):
        cfg = json.loads(json.dumps(base))
        if limited:
            cap = 0.5 * out[False][0]["sat"]["wells"][0]["last_surface_rates"][phase]
            cfg["wells"][0]["limits"] = {limit: cap}
        case = oi.Case(cfg, pvt=oi.PVT(PVT_SYNTH))
        sim = oi.Simulation(case)
        st = dict(sim.state)
        rec = sim.step(86400.0)
        out[limited] = (rec, st, case)
    free, lim = out[False][0]["sat"]["wells"][0], out[True][0]["sat"]["wells"][0]
    rec, st, case = out[True]
    p = rec["p"][2, 2, 0]
    kro, krw, krg = case.relperms(st["sw"][2, 2, 0], st["sg"][2, 2, 0])
    mu_o, mu_w, mu_g = (float(case.pvt(k, p)) * 1e-3 for k in ("mu_o", "mu_w", "mu_g"))
    mo, mw, mg = kro / mu_o, krw / mu_w, krg / mu_g
    mt = mo + mw + mg + 1e-10
    fo, fw, fg = mo / mt, mw / mt, mg / mt
    bo, bw, bg, rs, rv = (float(case.pvt(k, p)) for k in ("Bo", "Bw", "Bg", "Rs", "Rv"))
    q = abs(lim["last_q_total"])
    surf = {"oil": q * (fo / bo + fg * rv / bg) * 86400.0, "water": q * fw / bw * 86400.0,
            "gas": q * (fg / bg + fo * rs / bo) * 86400.0}
    surf["liquid"] = surf["oil"] + surf["water"]
    for k, v in surf.items():
        assert math.isclose(lim["last_surface_rates"][k], v, rel_tol=1e-12), k
    assert math.isclose(surf[phase], cap, rel_tol=1e-12)
    # the unlimited well produced its liquid control (no choke binds)
    assert math.isclose(free["last_surface_rates"]["liquid"], 8.0, rel_tol=1e-9)
