"""Tensor-level wrappers of the C ABI (torch tensors on a ROCm device in, out).

Index tensors are int32, value tensors float64 or float32, all contiguous and on
the same device. Launches go on the current torch stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import AmgParams, ImpesError, SolveStats, call, load, ptr, stream_ptr

__all__ = [
    "spmv_csr", "spmv_st7", "l1jacobi_csr", "dot", "AmgHierarchy", "FgmresSolver", "amg_default_params",
    "ktimer_reset", "ktimer_read", "copy_f64",
    "ImpesModel", "csr_diagonal",
]


def _i32(t):
    if t.dtype != torch.int32:
        raise ImpesError(f"expected int32 indices, got {t.dtype}")
    return t


def spmv_csr(rowptr, col, val, x, n_cols=None, out=None):
    n = rowptr.numel() - 1
    n_cols = x.numel() if n_cols is None else n_cols
    y = out if out is not None else torch.empty(n, dtype=val.dtype, device=val.device)
    fn = "impes_spmv_csr_f64" if val.dtype == torch.float64 else "impes_spmv_csr_f32"
    if x.dtype != val.dtype:
        raise ImpesError("x and val dtypes differ")
    call(fn, n, n_cols, val.numel(), ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val), ptr(x), ptr(y), stream_ptr())
    return y


def spmv_st7(dims, d, ox, oy, oz, x, out=None):
    nx, ny, nz = (int(v) for v in dims)
    y = out if out is not None else torch.empty_like(x)
    call("impes_spmv_st7_f64", nx, ny, nz, ptr(d), ptr(ox), ptr(oy), ptr(oz), ptr(x), ptr(y), stream_ptr())
    return y


def l1jacobi_csr(rowptr, col, val, inv_relax, omega, x_in, b, out=None):
    n = rowptr.numel() - 1
    y = out if out is not None else torch.empty_like(x_in)
    if val.dtype == torch.float64:
        call("impes_l1jacobi_csr_f64", n, val.numel(), ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val),
             ptr(inv_relax), float(omega), ptr(x_in), ptr(b), ptr(y), stream_ptr())
    else:
        call("impes_l1jacobi_csr_f32", n, val.numel(), ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val),
             ptr(inv_relax), C.c_float(omega), ptr(x_in), ptr(b), ptr(y), stream_ptr())
    return y


def dot(x, y):
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    call("impes_dot_f64", x.numel(), ptr(x), ptr(y), ptr(out), stream_ptr())
    return out


def csr_diagonal(rowptr, col, val):
    n = rowptr.numel() - 1
    out = torch.empty(n, dtype=torch.float64, device=val.device)
    call("impes_csr_diagonal", n, ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val), ptr(out), stream_ptr())
    return out


def copy_f64(y, x):
    """y = x with the library's streaming copy kernel (FP64, same length)."""
    assert x.numel() == y.numel() and x.dtype == y.dtype == torch.float64
    call("impes_copy_f64", x.numel(), ptr(x), ptr(y), stream_ptr())
    return y


_KT = {"name": ""}


def ktimer_reset(enable: bool = True, kernel: str = "st7_jacobi_f32"):
    """Select the kernel whose launches are timed with HIP event pairs (resets totals)."""
    _KT["name"] = kernel if enable else ""
    call("impes_ktimer_config", kernel.encode() if enable else None, 1 if enable else 0)


def ktimer_read() -> dict:
    """Totals since ktimer_reset: event-timed milliseconds, algorithmic bytes, launches."""
    ms, by, n = C.c_double(), C.c_double(), C.c_longlong()
    call("impes_ktimer_read", C.byref(ms), C.byref(by), C.byref(n))
    return {"name": _KT["name"], "ms": ms.value, "bytes": by.value, "launches": n.value}


def amg_default_params() -> AmgParams:
    p = AmgParams()
    load().impes_amg_default_params(C.byref(p))
    return p


class AmgHierarchy:
    """Owner of an ``impes_amg`` handle (device-resident RS-AMG hierarchy)."""

    # Reference behaviour of the four preconditioner choices that are not in the reference.
    REFERENCE = {"equil_mode": 0, "coarse_pin": 1, "trunc": 0.0, "pmax": 0, "coarsening": 0, "dominant": 0.0,
                 "pre_smooth": 3, "post_smooth": 2, "omega": 0.7}

    def __init__(self, theta=0.25, max_levels=10, coarsest_size=200, pre_smooth=2, post_smooth=2, omega=0.8,
                 equil_threshold=1e-6, dense_max=500, coarse_sweeps=50, anchor=None, equil_mode=1, coarse_pin=2,
                 trunc=0.2, pmax=4, coarsening=1, dominant=10.0):
        p = amg_default_params()
        p.coarsening = int(coarsening)
        p.dominant = float(dominant)
        p.equil_mode = int(equil_mode)
        p.coarse_pin = int(coarse_pin)
        p.trunc = float(trunc)
        p.pmax = int(pmax)
        p.theta = float(theta)
        p.max_levels = int(max_levels)
        p.coarsest_size = int(coarsest_size)
        p.pre_smooth = int(pre_smooth)
        p.post_smooth = int(post_smooth)
        p.omega = float(omega)
        p.equil_threshold = float(equil_threshold)
        p.dense_max = int(dense_max)
        p.coarse_sweeps = int(coarse_sweeps)
        p.anchor = -1 if anchor is None else int(anchor)
        self.params = p
        h = C.c_void_p()
        call("impes_amg_create", C.byref(p), C.byref(h))
        self._h = h
        self.n = 0
        self.device = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            try:
                load().impes_amg_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def setup_csr(self, rowptr, col, val):
        n = rowptr.numel() - 1
        call("impes_amg_setup_csr", self._h, n, val.numel(), ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val),
             stream_ptr())
        self.n = n
        self.device = val.device

    def setup_st7(self, dims, d, ox, oy, oz):
        nx, ny, nz = (int(v) for v in dims)
        call("impes_amg_setup_st7", self._h, nx, ny, nz, ptr(d), ptr(ox), ptr(oy), ptr(oz), stream_ptr())
        self.n = nx * ny * nz
        self.device = d.device

    def update_csr(self, val):
        call("impes_amg_update_csr", self._h, ptr(val), stream_ptr())

    def update_st7(self, d, ox, oy, oz):
        call("impes_amg_update_st7", self._h, ptr(d), ptr(ox), ptr(oy), ptr(oz), stream_ptr())

    @property
    def num_levels(self) -> int:
        v = C.c_int()
        call("impes_amg_num_levels", self._h, C.byref(v))
        return v.value

    def level_info(self, level: int) -> dict:
        n, nnz, pnnz, co, de = (C.c_int() for _ in range(5))
        call("impes_amg_level_info", self._h, int(level), C.byref(n), C.byref(nnz), C.byref(pnnz), C.byref(co),
             C.byref(de))
        return {"n": n.value, "nnz": nnz.value, "p_nnz": pnnz.value, "is_coarsest": bool(co.value),
                "dense": bool(de.value)}

    def level_times(self, precision=32, cycles=20):
        """Per-level V-cycle figures on the hierarchy's current V-cycle input
        (impes_amg_level_times): (ms per eager cycle of each level's own launches,
        their algorithmic bytes per visit, ms of one graph-replayed cycle)."""
        ms, by = (C.c_double * 64)(), (C.c_double * 64)()
        cyc = C.c_double()
        call("impes_amg_level_times", self._h, int(precision), int(cycles), ms, by, 64, C.byref(cyc), stream_ptr())
        nl = self.num_levels
        return [ms[i] for i in range(nl)], [by[i] for i in range(nl)], cyc.value

    @property
    def equilibrated(self) -> bool:
        v = C.c_int()
        call("impes_amg_equilibrated", self._h, C.byref(v))
        return bool(v.value)

    @property
    def setup_seconds(self) -> float:
        v = C.c_double()
        call("impes_amg_setup_seconds", self._h, C.byref(v))
        return v.value

    def export_level(self, level: int) -> dict:
        info = self.level_info(level)
        dev = self.device
        n, nnz, pnnz = info["n"], info["nnz"], info["p_nnz"]
        out = {
            "rowptr": torch.empty(n + 1, dtype=torch.int32, device=dev),
            "col": torch.empty(nnz, dtype=torch.int32, device=dev),
            "val": torch.empty(nnz, dtype=torch.float64, device=dev),
            "inv_relax": torch.empty(n, dtype=torch.float64, device=dev),
        }
        if not info["is_coarsest"]:
            out["p_rowptr"] = torch.empty(n + 1, dtype=torch.int32, device=dev)
            out["p_col"] = torch.empty(pnnz, dtype=torch.int32, device=dev)
            out["p_val"] = torch.empty(pnnz, dtype=torch.float64, device=dev)
            out["cf"] = torch.empty(n, dtype=torch.int8, device=dev)
        call("impes_amg_export_level", self._h, int(level), ptr(out["rowptr"]), ptr(out["col"]), ptr(out["val"]),
             ptr(out.get("p_rowptr")), ptr(out.get("p_col")), ptr(out.get("p_val")), ptr(out.get("cf")),
             ptr(out["inv_relax"]), stream_ptr())
        out.update(info)
        return out

    def dhalf_inv(self):
        t = torch.empty(self.n, dtype=torch.float64, device=self.device)
        call("impes_amg_export_scaling", self._h, ptr(t), stream_ptr())
        return t

    def vcycle(self, b, x=None, precision=64, transposed=False):
        """x <- V(b) in the equilibrated space (x=None: start from zero);
        transposed: the cycle with pre- and post-smoothing swapped (its transpose)."""
        zero = x is None
        if zero:
            x = torch.empty_like(b)
        fn = "impes_amg_vcycle_transposed" if transposed else "impes_amg_vcycle"
        call(fn, self._h, int(precision), ptr(b), ptr(x), 1 if zero else 0, stream_ptr())
        return x

    def set_graphs(self, enable: bool):
        """V-cycle replay from a captured HIP graph (default) or eager launches."""
        call("impes_amg_set_graphs", self._h, 1 if enable else 0)

    def precondition(self, v, precision=64, out=None):
        z = out if out is not None else torch.empty_like(v)
        call("impes_amg_precondition", self._h, int(precision), ptr(v), ptr(z), stream_ptr())
        return z

    def cycles(self, b, x=None, *, min_cycles=1, tol=1e-6, max_cycles=10, precision=64):
        """Stationary V-cycles in physical space (reference ClassicalAMG.solve/apply).
        Returns (x, relative residual, cycles done)."""
        use_x0 = x is not None
        out = x.clone() if use_x0 else torch.empty_like(b)
        rel = C.c_double()
        done = C.c_int()
        call("impes_amg_cycles", self._h, int(precision), ptr(b), ptr(out), 1 if use_x0 else 0, int(min_cycles),
             float(tol), int(max_cycles), C.byref(rel), C.byref(done), stream_ptr())
        return out, rel.value, done.value


class FgmresSolver:
    """Owner of an ``impes_solver`` handle (FGMRES workspace)."""

    def __init__(self):
        h = C.c_void_p()
        call("impes_solver_create", C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            try:
                load().impes_solver_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __del__(self):
        self.close()

    def solve_csr(self, amg, rowptr, col, val, dinv, b, x, *, precision=64, use_x0=False, rtol=1e-8,
                  restart=30, maxit=300, etol=0.0) -> dict:
        st = SolveStats()
        call("impes_solver_set_error_tol", self._h, float(etol))
        n = rowptr.numel() - 1
        call("impes_fgmres_csr", self._h, amg.handle if amg is not None else None, int(precision), n, val.numel(),
             ptr(_i32(rowptr)), ptr(_i32(col)), ptr(val), ptr(dinv), ptr(b), ptr(x), 1 if use_x0 else 0,
             float(rtol), int(restart), int(maxit), C.byref(st), stream_ptr())
        return st.as_dict()

    def solve_st7(self, amg, dims, d, ox, oy, oz, dinv, b, x, *, precision=64, use_x0=False, rtol=1e-8,
                  restart=30, maxit=300, etol=0.0) -> dict:
        st = SolveStats()
        call("impes_solver_set_error_tol", self._h, float(etol))
        nx, ny, nz = (int(v) for v in dims)
        call("impes_fgmres_st7", self._h, amg.handle if amg is not None else None, int(precision), nx, ny, nz,
             ptr(d), ptr(ox), ptr(oy), ptr(oz), ptr(dinv), ptr(b), ptr(x), 1 if use_x0 else 0, float(rtol),
             int(restart), int(maxit), C.byref(st), stream_ptr())
        return st.as_dict()


# ---------------------------------------------------------------------- IMPES model
REPORT_HEAD = 32
WELL_SLOTS = 16
EXPORT_FIELDS = {"diag": 0, "ox": 1, "oy": 2, "oz": 3, "rhs": 4, "acc": 5, "extra": 6, "tx": 7, "ty": 8,
                 "tz": 9, "pv": 10, "mob_t": 11, "dinv": 12}
PVT_KEYS = {"Bo": 0, "Bw": 1, "Bg": 2, "mu_o": 3, "mu_w": 4, "mu_g": 5, "Rs": 6, "Rv": 7}
PROP_NAMES = ("kro", "krw", "krg", "pc", "pcog", "ct", "rho_w", "rho_o", "rho_g", "mu_o", "mu_w", "mu_g")


class ImpesModel:
    """Owner of an ``impes_model`` handle: the device-resident IMPES step of one reservoir."""

    def __init__(self, grid: tuple, fluid: "FluidStruct", pvt_table=None, wells=()):
        from ._lib import FluidStruct, Grid, WellStruct  # noqa: F401

        nx, ny, nz, dx, dy, dz = grid
        g = Grid(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz))
        self.n = int(nx) * int(ny) * int(nz)
        self.dims = (int(nx), int(ny), int(nz))
        if pvt_table is not None:
            flat = (C.c_double * pvt_table.size)(*[float(v) for v in pvt_table.ravel()])
            pn = pvt_table.shape[1]
        else:
            flat, pn = None, 0
        warr, nw = self._wells(wells)
        h = C.c_void_p()
        call("impes_model_create", C.byref(g), C.byref(fluid), pn, flat, nw, warr, C.byref(h))
        self._h = h
        self.nwells = nw
        self._report = None

    @staticmethod
    def _wells(wells):
        from ._lib import WellStruct

        wells = list(wells)
        if not wells:
            return None, 0
        arr = (WellStruct * len(wells))(*wells)
        return arr, len(wells)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            try:
                load().impes_model_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __del__(self):
        self.close()

    def set_rock(self, kx, ky, kz, pv):
        call("impes_model_set_rock", self._h, ptr(kx), ptr(ky), ptr(kz), ptr(pv), stream_ptr())

    def set_envs(self, env_nx: int):
        """The grid holds nx // env_nx independent reservoirs stacked along x (call
        before set_rock); the saturation report then ends with their substep counts."""
        call("impes_model_set_envs", self._h, int(env_nx), stream_ptr())
        self._report = None

    def set_wells(self, wells):
        warr, nw = self._wells(wells)
        call("impes_model_set_wells", self._h, nw, warr, stream_ptr())
        self.nwells = nw
        self._report = None

    def set_amg(self, params: AmgParams):
        call("impes_model_set_amg", self._h, C.byref(params))

    def report_len(self) -> int:
        v = C.c_int()
        call("impes_model_report_len", self._h, C.byref(v))
        return v.value

    def assemble(self, dt, p, sw, sg):
        call("impes_model_assemble", self._h, float(dt), ptr(p), ptr(sw), ptr(sg), stream_ptr())

    def solve_pressure(self, p_prev, p_new, *, precision=32, rtol=1e-8, restart=30, maxit=300,
                       etol=0.0) -> dict:
        """FGMRES to the relative residual rtol; etol > 0 adds the pressure-error
        check (estimated relative error in the infinity norm held to etol)."""
        st = SolveStats()
        call("impes_model_set_error_tol", self._h, float(etol))
        call("impes_model_solve_pressure", self._h, int(precision), float(rtol), int(restart), int(maxit),
             ptr(p_prev), ptr(p_new), C.byref(st), stream_ptr())
        return st.as_dict()

    # ---- reverse sweep (hand-written adjoint of the step)
    def adjoint_reset(self):
        call("impes_model_adjoint_reset", self._h, stream_ptr())

    def transport_vjp(self, dt, p_prev, p_new, sw, sg, g_sw_next, g_sg_next, g_obs, g_p_new, g_p_prev, g_sw, g_sg):
        """g_obs: host array (nwells, 3). Writes g_sw, g_sg; adds into g_p_new, g_p_prev."""
        import numpy as np

        go = np.ascontiguousarray(np.asarray(g_obs, dtype=np.float64).reshape(-1))
        buf = (C.c_double * max(go.size, 1))(*go.tolist())
        call("impes_model_transport_vjp", self._h, float(dt), ptr(p_prev), ptr(p_new), ptr(sw), ptr(sg),
             ptr(g_sw_next), ptr(g_sg_next), buf, ptr(g_p_new), ptr(g_p_prev), ptr(g_sw), ptr(g_sg), stream_ptr())

    def solve_adjoint(self, g, lam, *, precision=32, rtol=1e-10, restart=30, maxit=400) -> dict:
        st = SolveStats()
        call("impes_model_solve_adjoint", self._h, int(precision), float(rtol), int(restart), int(maxit), ptr(g),
             ptr(lam), C.byref(st), stream_ptr())
        return st.as_dict()

    def assemble_vjp(self, dt, p_prev, sw, sg, p_new, lam, g_p_prev, g_sw, g_sg):
        call("impes_model_assemble_vjp", self._h, float(dt), ptr(p_prev), ptr(sw), ptr(sg), ptr(p_new), ptr(lam),
             ptr(g_p_prev), ptr(g_sw), ptr(g_sg), stream_ptr())

    def rock_vjp(self, kx, ky, kz):
        g = [torch.empty_like(kx) for _ in range(3)]
        call("impes_model_rock_vjp", self._h, ptr(kx), ptr(ky), ptr(kz), ptr(g[0]), ptr(g[1]), ptr(g[2]),
             stream_ptr())
        return g

    def adjoint_values(self):
        import numpy as np

        buf = (C.c_double * max(self.nwells, 1))()
        call("impes_model_adjoint_values", self._h, buf, stream_ptr())
        return np.array(buf[:self.nwells], dtype=np.float64)

    def solve_cg(self, p_new, *, tol, maxit) -> tuple[bool, int]:
        conv, its = C.c_int(), C.c_int()
        call("impes_model_solve_cg", self._h, float(tol), int(maxit), ptr(p_new), C.byref(conv), C.byref(its),
             stream_ptr())
        return bool(conv.value), its.value

    def saturation(self, dt, p_prev, p_new, sw, sg, sw_out, sg_out, so_out=None):
        """Runs the transport step; returns the raw report (numpy float64 array)."""
        import numpy as np

        n = self.report_len()
        if self._report is None or len(self._report) != n:
            self._report = (C.c_double * n)()
        call("impes_model_saturation", self._h, float(dt), ptr(p_prev), ptr(p_new), ptr(sw), ptr(sg), ptr(sw_out),
             ptr(sg_out), ptr(so_out), self._report, stream_ptr())
        return np.frombuffer(self._report, dtype=np.float64, count=n).copy()

    def export(self, field: str, out=None, device=None):
        t = out if out is not None else torch.empty(self.n, dtype=torch.float64, device=device or "cuda")
        call("impes_model_export", self._h, EXPORT_FIELDS[field], ptr(t), stream_ptr())
        return t

    def eval_props(self, p, sw, sg) -> dict:
        n = p.numel()
        out = torch.empty(12 * n, dtype=torch.float64, device=p.device)
        call("impes_model_eval_props", self._h, n, ptr(p), ptr(sw), ptr(sg), ptr(out), stream_ptr())
        return {k: out[i * n:(i + 1) * n] for i, k in enumerate(PROP_NAMES)}

    def eval_pvt(self, key: str, p):
        out = torch.empty_like(p)
        call("impes_model_eval_pvt", self._h, PVT_KEYS[key], p.numel(), ptr(p), ptr(out), stream_ptr())
        return out

    def amg_levels(self, max_levels=64):
        nl = C.c_int()
        sizes = (C.c_int * max_levels)()
        nnz = (C.c_int * max_levels)()
        call("impes_model_amg_levels", self._h, C.byref(nl), sizes, nnz, max_levels)
        return [(sizes[i], nnz[i]) for i in range(min(nl.value, max_levels))]

    def amg_level_times(self, precision=32, cycles=20):
        """Per-level V-cycle figures of the pressure hierarchy
        (impes_model_amg_level_times): (ms per eager cycle of each level's own
        launches, their algorithmic bytes per visit, ms of one graph-replayed cycle)."""
        ms, by = (C.c_double * 64)(), (C.c_double * 64)()
        cyc = C.c_double()
        call("impes_model_amg_level_times", self._h, int(precision), int(cycles), ms, by, 64, C.byref(cyc),
             stream_ptr())
        nl = len(self.amg_levels())
        return [ms[i] for i in range(nl)], [by[i] for i in range(nl)], cyc.value

    def amg_setup_seconds(self) -> float:
        """Wall time of the last hierarchy build (0.0 before the first solve)."""
        v = C.c_double()
        call("impes_model_amg_setup_seconds", self._h, C.byref(v))
        return v.value
