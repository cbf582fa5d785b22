"""Seeded synthetic cases of the benchmark suite (inputs, not physics).

``baseline_case(index, dims=None)`` returns ``(config, fields)``: a config
dict in the reference's JSON schema (reservoir / fluid / wells / simulation)
and per-cell fields in reservoir units (perm_x/y/z in mD, porosity) that
overwrite the reservoir's uniform fields. ``dims`` shrinks a case (wells are
placed proportionally) for parity tests.

  0  two-phase oil-water 60x60x1, lognormal k (median 100 mD, sigma_ln 0.5)
  1  three-phase 100x100x20, layered anisotropic correlated k, PVT, Stone II,
     capillarity, five-spot water injection, gravity
  2  three-phase gas injection 100x100x100 (1M cells), sigma_ln 1.5, 9 wells
  3  two-phase 64x64x8, correlated log-k (truth field of the history match),
     2 water injectors at 100 m^3/day and 3 liquid producers at 60 m^3/day
  4  = 2 (forward problem of the adjoint at scale)

The capillary pressure "Brooks-Corey, lambda = 2, entry 10 kPa" is realised
with the reference's power law pc_scale*(1 - s_n + 1e-6)^-e (fluid.py:273-287)
using pc_scale = entry pressure and e = 1/lambda.
"""
from __future__ import annotations

import numpy as np

SEEDS = {"field": 0, "medians": 1, "truth": 0, "prior": 1, "noise": 2, "fd": 3}

# Pressure-solver settings (the simulator's "amg" config) of the 1M-cell
# throughput benchmark: FP32 hierarchy inside FP64 FGMRES(30), started from the
# two-step extrapolation ("extrapolate2": the case's pressure alternates from step
# to step; 18.45 -> 18.15 iterations, 7.68 -> 7.62 ms per step at 1M cells against
# the linear one), stopped at a 1e-8 relative residual AND an estimated relative
# pressure error below 1e-8 (one preconditioner application per check);
# coarsening stops at 600 rows, where the coarsest level (471 rows at 1M cells) is
# inverted densely on the matrix cores at every update (5-7% faster per step than
# a ~100-row coarsest level; a 1254-row coarsest level runs about as fast). Strength threshold 0.1 (the reference's default is 0.25): these grids'
# vertical transmissibilities are ~10x the horizontal ones (kz = 0.1 kx, dz = dx / 10),
# so at 0.25 every horizontal coupling is weak and the split coarsens along z only,
# halving each level; at 0.1 it coarsens in all directions: operator complexity
# 3.06 -> 2.62, 19.6 -> 18.6 iterations, 11.6 -> 9.0 ms per step at 1M cells (0.08
# is as fast, 0.05 and 0.15 slower). The GPU parity tests run the shrunk benchmark
# cases with exactly these settings.
BENCH_AMG = {"tol": 1e-8, "error_tol": 1e-8, "precision": "fp32", "restart": 30, "initial_guess": "extrapolate2",
             "max_levels": 25, "max_iter": 400, "coarsest_size": 600, "dense_max": 600, "theta": 0.1}


def gaussian_field(shape, corr_cells: float, seed: int, axes=(0, 1)) -> np.ndarray:
    """Unit-variance Gaussian random field with a Gaussian covariance of length
    `corr_cells` along `axes` (FFT filtering of white noise), independent
    along the other axes."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    if corr_cells <= 0:
        return w
    f = np.fft.fftn(w, axes=axes)
    filt = np.ones([1] * len(shape))
    for ax in axes:
        k = np.fft.fftfreq(shape[ax])
        g = np.exp(-0.5 * (2 * np.pi * k * corr_cells) ** 2 / 2)
        sh = [1] * len(shape)
        sh[ax] = shape[ax]
        filt = filt * g.reshape(sh)
    out = np.real(np.fft.ifftn(f * filt, axes=axes))
    out -= out.mean()
    s = out.std()
    return out / s if s > 0 else out


def layered_logk(shape, lo_md, hi_md, sigma, corr, seed_field, seed_medians) -> np.ndarray:
    """ln k (mD): per layer (z) a log-uniform median in [lo, hi] plus a
    correlated Gaussian perturbation of standard deviation sigma."""
    nz = shape[2]
    rm = np.random.default_rng(seed_medians)
    med = np.exp(rm.uniform(np.log(lo_md), np.log(hi_md), size=nz))
    g = gaussian_field(shape, corr, seed_field, axes=(0, 1))
    return np.log(med)[None, None, :] + sigma * g


def porosity_from_logk(lnk, lo=0.1, hi=0.3) -> np.ndarray:
    a, b = lnk.min(), lnk.max()
    if b - a <= 0:
        return np.full(lnk.shape, 0.5 * (lo + hi))
    return lo + (hi - lo) * (lnk - a) / (b - a)


def black_oil_table(p_min=1.0, p_max=40.0, n=20, pb=18.0) -> dict:
    """20-point black-oil table in the reference's JSON schema (pvt.py:18-46):
    Bo, Rs and mu_o piecewise linear with a break at the bubble point pb,
    Bg = 0.1/p (p in MPa), mu_g = 0.02 cP, constant Bw and mu_w, no Rv."""
    P = np.linspace(p_min, p_max, n)
    rs = np.where(P <= pb, 10.0 + (120.0 - 10.0) * (P - p_min) / (pb - p_min), 120.0)
    bo = np.where(P <= pb, 1.10 + (1.35 - 1.10) * (P - p_min) / (pb - p_min), 1.35 - 0.002 * (P - pb))
    muo = np.where(P <= pb, 1.8 + (0.9 - 1.8) * (P - p_min) / (pb - p_min), 0.9 + 0.01 * (P - pb))
    return {"pressure_MPa": P.tolist(), "Bo": bo.tolist(), "Bw": [1.0] * n, "Bg": (0.1 / P).tolist(),
            "mu_o_cP": muo.tolist(), "mu_w_cP": [0.5] * n, "mu_g_cP": [0.02] * n, "Rs_m3m3": rs.tolist(),
            "Rv_m3m3": [0.0] * n, "units": {"pressure": "MPa"}, "meta": {"bubble_point_MPa": pb}}


def _place(ijk, full, dims):
    return [int(min(d - 1, round(c * d / f))) for c, f, d in zip(ijk, full, dims)]


def _well(name, kind, ijk, full, dims, **kw):
    i, j, k = _place(ijk, full, dims)
    return dict({"name": name, "type": kind, "i": i, "j": j, "k": k, "radius": 0.1}, **kw)


def baseline_case(index: int, dims=None, steps=None):
    index = 2 if index == 4 else index
    if index == 0:
        full = (60, 60, 1)
        dims = tuple(dims or full)
        lnk = np.log(100.0) + 0.5 * np.random.default_rng(SEEDS["field"]).standard_normal(dims)
        k = np.exp(lnk)
        fields = {"perm_x": k, "perm_y": k, "perm_z": k, "porosity": np.full(dims, 0.2)}
        cfg = {
            "reservoir": {"dimensions": list(dims), "grid_size": [10.0, 10.0, 5.0], "porosity": 0.2,
                          "permeability": 100.0, "k_vertical_fraction": 1.0},
            "fluid": {"pressure": 20.0, "s_w": 0.2, "s_g": 0.0, "mu_water": 1.0, "mu_oil": 5.0,
                      "c_oil": 7.5e-4, "c_water": 7.5e-4, "c_gas": 7.5e-4, "c_rock": 7.5e-4,
                      "relative_permeability": {"model": "corey", "nw": 2, "no": 2, "sw_cr": 0.2, "so_r": 0.2}},
            "wells": [_well("INJ", "injector", (0, 0, 0), full, dims, control_type="rate", control_value=100.0,
                            injected_phase="water"),
                      _well("PROD", "producer", (59, 59, 0), full, dims, control_type="bhp", control_value=15.0)],
            "simulation": {"solver_type": "impes", "time_step_days": 1.0, "steps": steps or 50,
                           "pressure_solver": "amg", "amg": dict(BENCH_AMG, precision="fp64")},
        }
        return cfg, fields
    if index in (1, 2):
        if index == 1:
            full = (100, 100, 20)
            sigma, grid, p0 = 1.0, [20.0, 20.0, 2.0], 20.0
        else:
            full = (100, 100, 100)
            sigma, grid, p0 = 1.5, [10.0, 10.0, 1.0], 25.0
        dims = tuple(dims or full)
        lnk = layered_logk(dims, 50.0, 500.0, sigma, 5.0, SEEDS["field"], SEEDS["medians"])
        kx = np.exp(lnk)
        fields = {"perm_x": kx, "perm_y": kx, "perm_z": 0.1 * kx, "porosity": porosity_from_logk(lnk)}
        fluid = {"pressure": p0, "s_w": 0.2, "s_g": 0.0, "pvt_table": black_oil_table(),
                 "relative_permeability": {"model": "stone2", "nw": 2, "no": 2, "ng": 2, "sw_cr": 0.2,
                                           "so_r": 0.2, "sg_cr": 0.0},
                 "capillary_pressure": {"pc_scale": 1e4, "pc_exponent": 0.5}}
        if index == 1:
            mid = full[2] // 2
            wells = [_well(f"INJ-{n}", "injector", c, full, dims, control_type="rate", control_value=200.0,
                           injected_phase="water")
                     for n, c in enumerate([(0, 0, mid), (0, 99, mid), (99, 0, mid), (99, 99, mid)])]
            wells.append(_well("PROD", "producer", (50, 50, mid), full, dims, control_type="bhp",
                               control_value=12.0))
            sim = {"time_step_days": 2.0, "steps": steps or 30}
        else:
            wells = [_well(f"GINJ-{n}", "injector", c, full, dims, control_type="rate", control_value=5e4,
                           rate_type="surface", injected_phase="gas")
                     for n, c in enumerate([(10, 10, 20), (10, 89, 20), (89, 10, 20), (89, 89, 20)])]
            wells += [_well(f"OPROD-{n}", "producer", c, full, dims, control_type="rate", control_value=300.0,
                            rate_type="surface", surface_phase="oil")
                      for n, c in enumerate([(30, 50, 80), (70, 50, 80), (50, 30, 80), (50, 70, 80)])]
            wells.append(_well("BHP-PROD", "producer", (50, 50, 50), full, dims, control_type="bhp",
                               control_value=10.0))
            sim = {"time_step_days": 1.0, "steps": steps or 20}
        sim.update({"solver_type": "impes", "pressure_solver": "amg", "amg": dict(BENCH_AMG)})
        cfg = {"reservoir": {"dimensions": list(dims), "grid_size": grid, "porosity": 0.2, "permeability": 100.0,
                             "k_vertical_fraction": 0.1},
               "fluid": fluid, "wells": wells, "simulation": sim}
        return cfg, fields
    if index == 3:
        full = (64, 64, 8)
        dims = tuple(dims or full)
        lnk = np.log(100.0) + 1.0 * gaussian_field(dims, 4.0, SEEDS["truth"], axes=(0, 1, 2))
        k = np.exp(lnk)
        fields = {"perm_x": k, "perm_y": k, "perm_z": k, "porosity": np.full(dims, 0.2)}
        cfg0, _ = baseline_case(0)
        cfg = {"reservoir": {"dimensions": list(dims), "grid_size": [10.0, 10.0, 5.0], "porosity": 0.2,
                             "permeability": 100.0, "k_vertical_fraction": 1.0},
               "fluid": cfg0["fluid"],
               "wells": [_well("INJ-1", "injector", (8, 8, 4), full, dims, control_type="rate",
                               control_value=100.0),
                         _well("INJ-2", "injector", (55, 55, 4), full, dims, control_type="rate",
                               control_value=100.0),
                         _well("PROD-1", "producer", (55, 8, 4), full, dims, control_type="rate",
                               control_value=60.0),
                         _well("PROD-2", "producer", (8, 55, 4), full, dims, control_type="rate",
                               control_value=60.0),
                         _well("PROD-3", "producer", (32, 32, 4), full, dims, control_type="rate",
                               control_value=60.0)],
               "simulation": {"solver_type": "impes", "time_step_days": 5.0, "steps": steps or 30,
                              "pressure_solver": "amg", "amg": dict(BENCH_AMG)}}
        return cfg, fields
    raise ValueError(f"no benchmark case {index}")


def build(config: dict, fields: dict | None = None, device="cuda"):
    """Reservoir, WellManager, Fluid and Simulator of a case (fields overwrite the
    uniform reservoir values before the Simulator is constructed)."""
    from .fluid import Fluid
    from .reservoir import Reservoir
    from .simulation import Simulator
    from .well import WellManager

    res = Reservoir(config["reservoir"], device)
    if fields:
        res.set_fields(fields.get("perm_x"), fields.get("perm_y"), fields.get("perm_z"), fields.get("porosity"))
    wm = WellManager(config["wells"], res)
    fl = Fluid(config["fluid"], res, device)
    sim = Simulator(res, fl, wm, config.get("simulation", {}), device)
    return sim


def rl_env_case() -> dict:
    """A small three-phase environment for the reinforcement-learning benchmark:
    20x20x3 cells of 10 x 10 x 5 m, gas and water rate injectors (50 m3/day each)
    and two BHP producers at 20 MPa, 1-day steps, pressure by FP32-AMG-preconditioned
    FGMRES to 1e-8."""
    return {
        "reservoir": {"dimensions": [20, 20, 3], "grid_size": [10.0, 10.0, 5.0], "porosity": 0.25,
                      "permeability": 150.0, "k_vertical_fraction": 0.1, "c_rock": 1e-5},
        "fluid": {"pressure": 25.0, "s_w": 0.2, "s_g": 0.05, "mu_oil": 2.0, "mu_water": 0.5, "mu_gas": 0.02,
                  "rho_oil": 850.0, "rho_water": 1000.0, "rho_gas": 1.2, "c_oil": 1e-5, "c_water": 5e-6,
                  "c_gas": 1e-3, "c_rock": 1e-5,
                  "relative_permeability": {"nw": 2.0, "no": 2.0, "ng": 2.5, "sw_cr": 0.15, "so_r": 0.2,
                                            "sg_cr": 0.02}},
        "wells": [
            {"name": "INJ-GAS", "type": "injector", "i": 2, "j": 10, "k": 1, "radius": 0.1,
             "control_type": "rate", "control_value": 50.0, "injected_phase": "gas"},
            {"name": "INJ-WATER", "type": "injector", "i": 17, "j": 10, "k": 1, "radius": 0.1,
             "control_type": "rate", "control_value": 50.0, "injected_phase": "water"},
            {"name": "PROD-1", "type": "producer", "i": 10, "j": 5, "k": 1, "radius": 0.1,
             "control_type": "bhp", "control_value": 20.0},
            {"name": "PROD-2", "type": "producer", "i": 10, "j": 15, "k": 1, "radius": 0.1,
             "control_type": "bhp", "control_value": 20.0},
        ],
        "simulation": {"solver_type": "impes", "time_step_days": 1.0, "max_time_step_attempts": 3,
                       "dt_reduction_factor": 2.0, "max_saturation_change": 0.05, "pressure_solver": "amg",
                       "amg": {"tol": 1e-8, "max_iter": 400}},
    }
