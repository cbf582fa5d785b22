"""Black-oil fluid model (reference: src/simulator/fluid.py:10-510).

Same configuration keys, defaults and public methods as the reference. The
state (pressure, s_w, s_g, s_o) is float64 on the ROCm device. Property
methods evaluate on the device through the HIP library (the same device
functions the IMPES kernels use); there is no host fallback.
The derivative helpers of the reference's fully-implicit solver are not part
of the IMPES path and are not provided.
"""
from __future__ import annotations

import torch

from .pvt import PVTTable

class Fluid:
    def __init__(self, config, reservoir, device=None):
        self.device = torch.device(device) if device is not None else reservoir.device
        self.dimensions = list(reservoir.dimensions)
        initial_pressure = float(config.get("pressure", 20.0)) * 1e6
        initial_sw = float(config.get("s_w", 0.2))
        initial_sg = float(config.get("s_g", 0.0))
        self.mu_oil = float(config.get("mu_oil", 1.0)) * 1e-3
        self.mu_water = float(config.get("mu_water", 0.5)) * 1e-3
        self.mu_gas = float(config.get("mu_gas", 0.02)) * 1e-3
        self.rho_oil_ref = float(config.get("rho_oil", 850.0))
        self.rho_water_ref = float(config.get("rho_water", 1000.0))
        self.rho_gas_ref = float(config.get("rho_gas", 1.0))
        self.oil_compressibility = float(config.get("c_oil", 1e-5)) / 1e6
        self.water_compressibility = float(config.get("c_water", 1e-5)) / 1e6
        self.gas_compressibility = float(config.get("c_gas", 1e-3)) / 1e6
        self.rock_compressibility = float(config.get("c_rock", 1e-5)) / 1e6
        total_c = (self.oil_compressibility + self.water_compressibility + self.gas_compressibility
                   + self.rock_compressibility) / 3
        f64 = dict(dtype=torch.float64, device=self.device)
        self.cf = torch.full(self.dimensions, total_c, **f64)
        self.pvt = None
        pvt_path = config.get("pvt_path") or (config.get("pvt", {}) or {}).get("path")
        if config.get("pvt_table") is not None:
            self.pvt = PVTTable(data=config["pvt_table"])
        elif pvt_path:
            try:
                self.pvt = PVTTable(pvt_path)
            except Exception as e:  # the reference keeps running without the table (fluid.py:54-58)
                print(f"  PVT: could not load '{pvt_path}': {e}")
        rp = config.get("relative_permeability", {})
        self.relperm_model = rp.get("model", "corey")
        self.nw = rp.get("nw", 2)
        self.no = rp.get("no", 2)
        self.ng = rp.get("ng", 2)
        self.ko_end_w = rp.get("ko_end_w", 1.0)
        self.ko_end_g = rp.get("ko_end_g", 1.0)
        self.now = rp.get("now", float(self.no))
        self.nog = rp.get("nog", float(self.ng))
        self.krw_end = rp.get("krw_end", 1.0)
        self.krg_end = rp.get("krg_end", 1.0)
        self.sw_cr = rp.get("sw_cr", 0.2)
        self.so_r = rp.get("so_r", 0.2)
        self.sg_cr = rp.get("sg_cr", 0.0)
        self.pressure = torch.full(self.dimensions, initial_pressure, **f64)
        self.s_w = torch.full(self.dimensions, initial_sw, **f64)
        self.s_g = torch.full(self.dimensions, initial_sg, **f64)
        self.s_o = torch.full(self.dimensions, 1.0 - initial_sw - initial_sg, **f64)
        if initial_sw + initial_sg > 1.0:
            raise ValueError("initial water + gas saturation exceeds 1.0")
        self.prev_pressure = self.pressure
        self.prev_sw = self.s_w
        self.prev_sg = self.s_g
        if "capillary_pressure" in config:
            pc = config["capillary_pressure"]
            self.pc_scale = pc.get("pc_scale", 0.0)
            self.pc_exponent = pc.get("pc_exponent", 1.5)
            self.pc_threshold = pc.get("pc_threshold", 0.01)
        else:
            self.pc_scale = config.get("pc_scale", 0.0)
            self.pc_exponent = config.get("pc_exponent", 1.5)
            self.pc_threshold = config.get("pc_threshold", 0.01)
        pcg = config.get("capillary_pressure_og", {})
        self.pc_og_scale = pcg.get("pc_scale", 0.0)
        self.pc_og_exponent = pcg.get("pc_exponent", 1.5)
        self._props_model = None

    # ------------------------------------------------------------ device parameters
    def fluid_struct(self, sim_params: dict | None = None):
        """Parameters of the HIP kernels (impes_fluid)."""
        from impes_hip import FluidStruct

        sp = sim_params or {}
        f = FluidStruct()
        f.stone2 = 1 if self.relperm_model == "stone2" else 0
        f.nw, f.no, f.ng = float(self.nw), float(self.no), float(self.ng)
        f.ko_end_w, f.ko_end_g = float(self.ko_end_w), float(self.ko_end_g)
        f.now_, f.nog = float(self.now), float(self.nog)
        f.sw_cr, f.so_r, f.sg_cr = float(self.sw_cr), float(self.so_r), float(self.sg_cr)
        f.mu_o, f.mu_w, f.mu_g = self.mu_oil, self.mu_water, self.mu_gas
        f.rho_o_ref, f.rho_w_ref, f.rho_g_ref = self.rho_oil_ref, self.rho_water_ref, self.rho_gas_ref
        f.c_o, f.c_w, f.c_g, f.c_r = (self.oil_compressibility, self.water_compressibility,
                                      self.gas_compressibility, self.rock_compressibility)
        f.cf = (f.c_o + f.c_w + f.c_g + f.c_r) / 3
        f.pc_scale, f.pc_exp = float(self.pc_scale), float(self.pc_exponent)
        f.pcog_scale, f.pcog_exp = float(self.pc_og_scale), float(self.pc_og_exponent)
        f.g = 9.81
        f.max_ds = float(sp.get("max_saturation_change", 0.05))
        f.max_substeps = int(sp.get("max_substeps", 20))
        f.rate_scale = float(sp.get("global_rate_scale", 1.0))
        return f

    def pvt_device_table(self):
        return None if self.pvt is None else self.pvt.device_table()

    def _model(self):
        if self._props_model is None:
            from impes_hip import ImpesModel

            self._props_model = ImpesModel((1, 1, 1, 1.0, 1.0, 1.0), self.fluid_struct(), self.pvt_device_table())
        return self._props_model

    def _eval(self, p=None, sw=None, sg=None) -> dict:
        ref = next(t for t in (sw, sg, p) if t is not None)
        ref = torch.as_tensor(ref, device=self.device)
        shape = ref.shape

        def prep(t, default):
            if t is None:
                t = default
            t = torch.as_tensor(t, dtype=torch.float64, device=self.device)
            return torch.broadcast_to(t, shape).contiguous().reshape(-1)

        zero = torch.zeros((), dtype=torch.float64, device=self.device)
        out = self._model().eval_props(prep(p, zero + 2e7), prep(sw, zero + 0.2), prep(sg, zero))
        return {k: v.reshape(shape) for k, v in out.items()}

    # ------------------------------------------------------------ reference API
    @property
    def rho_w(self):
        return self._eval(p=self.pressure, sw=self.s_w, sg=self.s_g)["rho_w"]

    @property
    def rho_o(self):
        return self._eval(p=self.pressure, sw=self.s_w, sg=self.s_g)["rho_o"]

    @property
    def rho_g(self):
        return self._eval(p=self.pressure, sw=self.s_w, sg=self.s_g)["rho_g"]

    @property
    def mu_w(self):
        return self.mu_water if self.pvt is None else self._eval_pvt(self.pressure, "mu_w") * 1e-3

    @property
    def mu_o(self):
        return self.mu_oil if self.pvt is None else self._eval_pvt(self.pressure, "mu_o") * 1e-3

    @property
    def mu_g(self):
        return self.mu_gas if self.pvt is None else self._eval_pvt(self.pressure, "mu_g") * 1e-3

    def _eval_pvt(self, p_tensor, key: str):
        if self.pvt is None:
            raise RuntimeError("PVT table not loaded")
        p = torch.as_tensor(p_tensor, dtype=torch.float64, device=self.device)
        return self._model().eval_pvt(key, p.contiguous().reshape(-1)).reshape(p.shape)

    def calc_total_compressibility(self, pressure, s_w, s_g):
        if self.pvt is None:
            return self.cf
        return self._eval(p=pressure, sw=s_w, sg=s_g)["ct"]

    def _get_normalized_saturation(self, s_w):
        s = torch.as_tensor(s_w, dtype=torch.float64, device=self.device)
        return torch.clamp((s - self.sw_cr) / (1 - self.sw_cr - self.so_r), 0.0, 1.0)

    def get_rel_perms(self, s_w, s_g=None):
        r = self._eval(sw=s_w, sg=self.s_g if s_g is None else s_g)
        return r["kro"], r["krw"], r["krg"]

    def calc_water_kr(self, s_w):
        return self._eval(sw=s_w)["krw"]

    def calc_oil_kr(self, s_w, s_g=None):
        return self._eval(sw=s_w, sg=self.s_g if s_g is None else s_g)["kro"]

    def calc_gas_kr(self, s_g):
        return self._eval(sg=s_g)["krg"]

    def get_capillary_pressure(self, s_w):
        return self._eval(sw=s_w)["pc"]

    def get_capillary_pressure_og(self, s_g):
        return self._eval(sg=s_g)["pcog"]

    calc_capillary_pressure = get_capillary_pressure

    def calc_water_density(self, pressure):
        return self._eval(p=pressure)["rho_w"]

    def calc_oil_density(self, pressure):
        return self._eval(p=pressure)["rho_o"]

    def calc_gas_density(self, pressure):
        return self._eval(p=pressure)["rho_g"]
