"""IMPES simulator (reference: src/simulator/simulation.py:19-135, 822-1697).

``Simulator(reservoir, fluid, well_manager, sim_params, device).run_step(dt)``
keeps the reference's constructor, step entry, state fields
(fluid.pressure/s_w/s_g/s_o), ledgers (mass_balance, component_balance) and
per-well diagnostics (last_q_total, last_surface_rates, last_mode). Every
numeric stage runs in the HIP library on the ROCm device:

  pressure assembly (impes_model_assemble)
  pressure solve: "amg" / "gmres_amg" -> FP64 FGMRES with an RS-AMG V-cycle
                  preconditioner (FP32 hierarchy by default, sim_params["amg"]["precision"]);
                  "cg" (reference default) -> Jacobi-PCG with the reference's absolute stop
  transport      (impes_model_saturation)

Differences from the reference, by design: state is float64; the AMG branch
returns the solver's real convergence flag (the reference hard-codes True,
simulation.py:927) so the dt-reduction retry also covers it; the rate of a
BHP well in the transport step uses the pressure equation's row balance
(equal in exact arithmetic to WI*lambda*(p_cell - p_bhp), which cancels
catastrophically); a new volume ledger ``volume_balance`` closes
sum PV dS = dt sum q + clamp corrections per phase.
"""
from __future__ import annotations

import numpy as np
import torch

from .well import LIMIT_KEYS, SURFACE_PHASES

REPORT_HEAD = 32
WELL_SLOTS = 16


def _flat(t: torch.Tensor) -> torch.Tensor:
    t = t if t.dtype == torch.float64 else t.double()
    return t.contiguous().reshape(-1)


class Simulator:
    def __init__(self, reservoir, fluid, well_manager, sim_params, device=None, *, env_nx=None):
        """env_nx (not in the reference): the grid holds nx // env_nx independent
        reservoirs stacked along x (batched environments, rl_env.BatchedReservoirEnv):
        no flow between them and a substep count each (``last_substeps_env``)."""
        import impes_hip

        self.reservoir = reservoir
        self.fluid = fluid
        self.well_manager = well_manager
        self.sim_params = sim_params
        self.device = torch.device(device) if device is not None else fluid.device
        if self.device.type != "cuda":
            raise RuntimeError("this simulator runs on a ROCm GPU: pass device='cuda'")
        self.solver_type = sim_params.get("solver_type", "impes")
        if self.solver_type != "impes":
            raise NotImplementedError(f"solver_type '{self.solver_type}': only the IMPES scheme is provided")
        self.verbose = bool(sim_params.get("verbose", False))
        nx, ny, nz = (int(v) for v in reservoir.dimensions)
        self.dims = (nx, ny, nz)
        dx, dy, dz = float(reservoir.dx), float(reservoir.dy), float(reservoir.dz)
        self._model = impes_hip.ImpesModel((nx, ny, nz, dx, dy, dz), fluid.fluid_struct(sim_params),
                                           fluid.pvt_device_table(), self._well_structs())
        self.num_envs = 1
        if env_nx:
            if nx % int(env_nx):
                raise ValueError(f"nx={nx} is not a multiple of env_nx={env_nx}")
            self._model.set_envs(int(env_nx))
            self.num_envs = nx // int(env_nx)
        self.last_substeps_env = None
        kx, ky, kz = (_flat(t) for t in reservoir.permeability_tensors)
        self.porous_volume = reservoir.porous_volume
        self._model.set_rock(kx, ky, kz, _flat(self.porous_volume))
        self._amg_cfg = dict(sim_params.get("amg", {}) or {})
        p = impes_hip.amg_default_params()
        p.theta = float(self._amg_cfg.get("theta", 0.25))
        p.max_levels = int(self._amg_cfg.get("max_levels", 10))
        p.coarsest_size = int(self._amg_cfg.get("coarsest_size", 200))
        for key, typ in (("equil_mode", int), ("coarse_pin", int), ("trunc", float), ("pmax", int),
                         ("coarsening", int), ("dominant", float), ("pre_smooth", int), ("post_smooth", int),
                         ("omega", float), ("dense_max", int)):
            if key in self._amg_cfg:
                setattr(p, key, typ(self._amg_cfg[key]))
        self._model.set_amg(p)
        tx = self._model.export("tx", device=self.device).view(nx, ny, nz)
        ty = self._model.export("ty", device=self.device).view(nx, ny, nz)
        tz = self._model.export("tz", device=self.device).view(nx, ny, nz)
        self.T_x, self.T_y, self.T_z = tx[:-1], ty[:, :-1], tz[:, :, :-1]
        self.g = 9.81
        self.mass_balance = {ph: {"in": 0.0, "out": 0.0} for ph in ("water", "oil", "gas")}
        self.component_balance = {ph: {"in": 0.0, "out": 0.0, "accum": 0.0} for ph in ("oil", "gas")}
        self.volume_balance = {k: 0.0 for k in ("dV_w", "dV_g", "q_w", "q_g", "clamp_w", "clamp_g")}
        self.last_solve = {}
        self.last_substeps = None
        self.step_log = []
        self._p_prev = None
        self._p_hist = None  # (pressure at the start of the last accepted step, its dt)
        self._p_hist2 = None  # the same for the accepted step before it

    # ------------------------------------------------------------ wells
    def _ijk_to_idx(self, i, j, k, nx=None, ny=None):
        """simulation.py:804-816: z-fastest flat index."""
        _, ny_, nz_ = self.dims
        return (i * ny_ + j) * nz_ + k

    def _well_structs(self):
        from impes_hip import WellStruct

        nx, ny, nz = (int(v) for v in self.reservoir.dimensions)
        out = []
        for w in self.well_manager.get_wells():
            if not (0 <= w.i < nx and 0 <= w.j < ny and 0 <= w.k < nz):
                raise ValueError(f"well {w.name} at ({w.i},{w.j},{w.k}) lies outside the grid")
            s = WellStruct()
            s.cell = (w.i * ny + w.j) * nz + w.k
            w.cell_index_flat = s.cell  # as _calculate_well_terms does (simulation.py:1629-1630)
            s.producer = 0 if w.type == "injector" else 1
            s.control = {"rate": 0, "bhp": 1}.get(w.control_type, 2)
            s.gas = 1 if getattr(w, "injected_phase", "water") == "gas" else 0
            s.surface = 1 if getattr(w, "rate_type", "reservoir") == "surface" else 0
            s.surface_phase = SURFACE_PHASES.get((w.surface_phase or "liquid").lower(), 3)
            s.has_bhp_min = 1 if w.bhp_min is not None else 0
            s.bhp_min = float(w.bhp_min) if w.bhp_min is not None else 0.0
            s.value = float(w.control_value)
            s.wi = float(w.well_index)
            lim = getattr(w, "limits", None) or {}
            mask = 0
            for b, key in enumerate(LIMIT_KEYS):
                if key in lim:
                    mask |= 1 << b
                    s.lim[b] = float(lim[key])
            s.lim_mask = mask
            out.append(s)
        return out

    def update_wells(self):
        """Push changed well controls (e.g. from an RL policy) to the device."""
        self._model.set_wells(self._well_structs())

    # ------------------------------------------------------------ step
    def run_step(self, dt):
        if self.solver_type == "impes":
            return self._impes_step(dt)
        raise ValueError(f"unknown solver type: {self.solver_type}")

    def _impes_step(self, dt):
        """simulation.py:822-865: attempt loop with dt reduction."""
        original_dt = dt
        current_dt = dt
        max_attempts = self.sim_params.get("max_time_step_attempts", 5)
        dt_reduction_factor = self.sim_params.get("dt_reduction_factor", 2.0)
        for _attempt in range(max_attempts):
            P_new, converged = self._impes_pressure_step(current_dt)
            if converged:
                self._p_hist2 = self._p_hist
                self._p_hist = (self._p_prev.clone(), current_dt)
                self.fluid.pressure = P_new
                self._impes_saturation_step(P_new, current_dt)
                self.fluid.prev_pressure = self.fluid.pressure
                self.fluid.prev_sw = self.fluid.s_w
                self.last_dt = current_dt
                return True
            if self.verbose:
                print("  IMPES pressure solve did not converge: reducing dt")
            current_dt /= dt_reduction_factor
        self.last_dt = original_dt
        return False

    def _impes_pressure_step(self, dt):
        """simulation.py:867-966 on the device. Returns (P_new, converged)."""
        fl = self.fluid
        p, sw, sg = _flat(fl.pressure), _flat(fl.s_w), _flat(fl.s_g)
        self._p_prev = p
        self._model.assemble(dt, p, sw, sg)
        solver = self.sim_params.get("pressure_solver", "cg")
        p_new = torch.empty_like(p)
        if solver in ("amg", "gmres_amg", "fgmres_amg", "cg_amg"):
            cfg = self._amg_cfg
            tol = float(cfg.get("tol", self.sim_params.get("cg_tolerance", 1e-6)))
            prec = str(cfg.get("precision", "fp32")).lower()
            precision = 64 if prec in ("fp64", "64", "float64", "double") else 32
            # "error_tol": estimated relative pressure error (infinity norm) the solve
            # must also reach; defaults to tol, 0 stops on the residual alone
            etol = float(cfg.get("error_tol", tol))
            st = self._model.solve_pressure(self._initial_guess(p, dt, cfg), p_new, precision=precision, rtol=tol,
                                            restart=int(cfg.get("restart", 30)), maxit=int(cfg.get("max_iter", 300)),
                                            etol=etol)
            converged = bool(st["converged"])
        else:
            tol = float(self.sim_params.get("cg_tolerance", 1e-6))
            maxit = int(self.sim_params.get("cg_max_iter", 500))
            converged, its = self._model.solve_cg(p_new, tol=tol, maxit=maxit)
            if not converged:
                converged, its2 = self._model.solve_cg(p_new, tol=tol * 10.0, maxit=maxit * 4)
                its += its2
            st = {"iterations": its, "converged": int(converged), "solver": "cg"}
        self.last_solve = st
        return p_new.view(self.dims), converged

    def _initial_guess(self, p, dt, cfg):
        """FGMRES starting pressure. "previous": the current pressure. "extrapolate"
        (default): linear in time through the start pressures of this and the last
        accepted step, p + (p - p_last) dt / dt_last, when the last step was at
        least as long as this one; after a shorter step (one the attempt loop cut)
        the current pressure, since the change over a cut step says little about
        the next full one. "extrapolate2": with three equal steps behind it,
        p_last + (p - p_last2), the change over two steps added to the pressure
        two steps back, which is exact for a linear trend plus a component that
        alternates from step to step (as the pressure of cases with unconverted
        surface-rate injectors does); otherwise as "extrapolate". Either way the
        solve runs to the same stopping test; a better start saves iterations."""
        mode = str(cfg.get("initial_guess", "extrapolate"))
        if mode == "zero":
            return torch.zeros_like(p)
        if mode not in ("extrapolate", "extrapolate2") or self._p_hist is None:
            return p
        p_last, dt_last = self._p_hist
        if p_last.shape != p.shape or not float(dt_last) >= float(dt):
            return p
        if mode == "extrapolate2" and self._p_hist2 is not None:
            p_last2, dt_last2 = self._p_hist2
            if p_last2.shape == p.shape and float(dt_last) == float(dt) and float(dt_last2) == float(dt):
                return p_last + (p - p_last2)
        return torch.add(p, p - p_last, alpha=float(dt) / float(dt_last))

    def _impes_saturation_step(self, P_new, dt):
        """simulation.py:968-1544 on the device; ledgers and diagnostics from the report."""
        fl = self.fluid
        sw, sg = _flat(fl.s_w), _flat(fl.s_g)
        p_prev = self._p_prev if self._p_prev is not None else _flat(P_new)
        sw_out, sg_out, so_out = torch.empty_like(sw), torch.empty_like(sg), torch.empty_like(sw)
        rep = self._model.saturation(dt, p_prev, _flat(P_new), sw, sg, sw_out, sg_out, so_out)
        fl.s_w, fl.s_g, fl.s_o = sw_out.view(self.dims), sg_out.view(self.dims), so_out.view(self.dims)
        mb = self.mass_balance
        mb["water"]["in"] += rep[2]; mb["water"]["out"] += rep[3]
        mb["oil"]["in"] += rep[4]; mb["oil"]["out"] += rep[5]
        mb["gas"]["in"] += rep[6]; mb["gas"]["out"] += rep[7]
        cb = self.component_balance
        cb["oil"]["in"] += rep[8]; cb["oil"]["out"] += rep[9]
        cb["gas"]["in"] += rep[10]; cb["gas"]["out"] += rep[11]
        cb["oil"]["accum"] = cb["oil"]["in"] - cb["oil"]["out"]
        cb["gas"]["accum"] = cb["gas"]["in"] - cb["gas"]["out"]
        vb = self.volume_balance
        for i, k in enumerate(("dV_w", "dV_g", "q_w", "q_g", "clamp_w", "clamp_g")):
            vb[k] += rep[12 + i]
        self.last_substeps = int(rep[1])
        if self.num_envs > 1:
            self.last_substeps_env = [int(v) for v in rep[len(rep) - self.num_envs:]]
        for w, well in enumerate(self.well_manager.get_wells()):
            R = rep[REPORT_HEAD + WELL_SLOTS * w: REPORT_HEAD + WELL_SLOTS * (w + 1)]
            if R[0] == 0.0:
                well.last_mode = "rate"
            elif R[0] == 1.0:
                well.last_mode = "bhp"
            if R[3] == 1.0:
                well.last_q_total = float(R[4])
                well.last_surface_rates = {"oil": float(R[5]), "water": float(R[6]), "gas": float(R[7]),
                                           "liquid": float(R[8])}
        self.step_log.append({"dt": dt, "substeps": self.last_substeps, "rat_max": float(rep[0]),
                              "solve": dict(self.last_solve)})
        if self.verbose:
            print(f"P mean = {float(P_new.mean()) / 1e6:.2f} MPa, substeps N={self.last_substeps}", flush=True)

    # ------------------------------------------------------------ run
    def run(self, output_filename, save_vtk=False, results_root="results"):
        """simulation.py:2120-2248: total_time_days / time_step_days steps, then the
        run's artifacts in results_root/<name>_<timestamp>/: <name>.txt (the same
        SUMMARY and MASS BALANCE lines; full arrays when write_full_arrays_txt) and
        <name>.npz (pressure, sw, sg, so; when save_npz, default True). Every
        save_interval steps the fields go to intermediate/<name>_step_<i>.npz
        (the reference renders PNG plots there; plotting is not part of this
        package, nor is VTK output). Returns the results directory."""
        import datetime
        import os

        import numpy as np

        total_days = self.sim_params.get("total_time_days", 100)
        step_days = self.sim_params.get("time_step_days", 1.0)
        save_interval = self.sim_params.get("save_interval", 10)
        num_steps = int(total_days / step_days)
        dt = step_days * 86400
        stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
        results_dir = os.path.join(results_root, f"{output_filename}_{stamp}")
        os.makedirs(results_dir, exist_ok=True)
        inter = os.path.join(results_dir, "intermediate")
        if save_interval < num_steps:
            os.makedirs(inter, exist_ok=True)
        fl = self.fluid

        def fields():
            return (fl.pressure.double().cpu().numpy(), fl.s_w.double().cpu().numpy(),
                    fl.s_g.double().cpu().numpy(), fl.s_o.double().cpu().numpy())

        for i in range(num_steps):
            ok = self.run_step(dt)
            if not ok and self.verbose:
                print(f"  step {i + 1}: pressure solve did not converge")
            if (i + 1) % save_interval == 0 and save_interval < num_steps:
                p, sw, sg, so = fields()
                np.savez_compressed(os.path.join(inter, f"{output_filename}_step_{i + 1}.npz"), pressure=p, sw=sw,
                                    sg=sg, so=so, day=(i + 1) * step_days)
        p, sw, sg, so = fields()
        with open(os.path.join(results_dir, f"{output_filename}.txt"), "w") as f:
            f.write("=== SUMMARY (final) ===\n")
            f.write(f"Pressure MPa: min={p.min()/1e6:.2f}, max={p.max()/1e6:.2f}, mean={p.mean()/1e6:.2f}\n")
            f.write(f"Sw: min={sw.min():.4f}, max={sw.max():.4f}, mean={sw.mean():.4f}\n")
            f.write(f"Sg: min={sg.min():.4f}, max={sg.max():.4f}, mean={sg.mean():.4f}\n")
            f.write(f"So: min={so.min():.4f}, max={so.max():.4f}, mean={so.mean():.4f}\n")
            mb = self.mass_balance
            net = {ph: mb[ph]["in"] - mb[ph]["out"] for ph in ("water", "gas", "oil")}
            f.write("\n=== MASS BALANCE (kg) ===\n")
            f.write(f"water: in={mb['water']['in']:.6e}, out={mb['water']['out']:.6e}, net={net['water']:.6e}\n")
            f.write(f"gas:   in={mb['gas']['in']:.6e}, out={mb['gas']['out']:.6e}, net={net['gas']:.6e}\n")
            f.write(f"oil:   in={mb['oil']['in']:.6e}, out={mb['oil']['out']:.6e}, net={net['oil']:.6e}\n")
            f.write(f"total net={net['water'] + net['gas'] + net['oil']:.6e}\n")
            if self.sim_params.get("write_full_arrays_txt", False):
                for title, arr, fmt in (("Final Pressure (MPa)", p / 1e6, "%.2f"), ("Final Water Saturation", sw, "%.4f"),
                                        ("Final Gas Saturation", sg, "%.4f"), ("Final Oil Saturation", so, "%.4f")):
                    f.write(f"\n\n{title}:\n")
                    f.write(np.array2string(arr, threshold=np.inf, formatter={"float_kind": lambda x, fm=fmt: fm % x}))
        if self.sim_params.get("save_npz", True):
            np.savez_compressed(os.path.join(results_dir, f"{output_filename}.npz"), pressure=p, sw=sw, sg=sg, so=so)
        if save_vtk:
            print("VTK output is not part of this package")
        return results_dir

    # ------------------------------------------------------------ diagnostics
    def volume_balance_error(self) -> dict:
        """Cumulative |sum PV dS - dt sum q - clamp| per phase over the total pore volume."""
        v = self.volume_balance
        pv = float(self.porous_volume.double().sum())
        return {"water": abs(v["dV_w"] - v["q_w"] - v["clamp_w"]) / pv,
                "gas": abs(v["dV_g"] - v["q_g"] - v["clamp_g"]) / pv}

    def pressure_system(self) -> dict:
        """The last assembled pressure system in face form (device tensors)."""
        names = ("diag", "ox", "oy", "oz", "rhs", "acc", "extra", "dinv", "mob_t")
        return {k: self._model.export(k, device=self.device) for k in names}

    def amg_levels(self):
        return self._model.amg_levels()
