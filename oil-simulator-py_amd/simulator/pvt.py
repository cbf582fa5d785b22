"""PVT table of the black-oil model (reference: src/simulator/pvt.py:18-64).

Same JSON schema and constructor checks as the reference. The table is
evaluated on the device during the IMPES step (impes_model_eval_pvt /
the transport kernels); ``eval`` is the host-side lookup of the reference API
(np.interp clamped at the ends) for callers that hold numpy arrays.
"""
from __future__ import annotations

import json

import numpy as np

COLUMNS = ("pressure_MPa", "Bo", "Bw", "Bg", "mu_o_cP", "mu_w_cP", "mu_g_cP", "Rs_m3m3", "Rv_m3m3")


class PVTTable:
    def __init__(self, json_path: str | None = None, data: dict | None = None):
        if data is None:
            with open(json_path, "r", encoding="utf-8") as f:
                data = json.load(f)
        self.P = np.asarray(data["pressure_MPa"], dtype=np.float64)
        self.Bo = np.asarray(data.get("Bo", []), dtype=np.float64)
        self.Bw = np.asarray(data.get("Bw", []), dtype=np.float64)
        self.Bg = np.asarray(data.get("Bg", []), dtype=np.float64)
        self.mu_o = np.asarray(data.get("mu_o_cP", []), dtype=np.float64)
        self.mu_w = np.asarray(data.get("mu_w_cP", []), dtype=np.float64)
        self.mu_g = np.asarray(data.get("mu_g_cP", []), dtype=np.float64)
        self.Rs = np.asarray(data.get("Rs_m3m3", []), dtype=np.float64)
        rv = data.get("Rv_m3m3")
        self.Rv = np.zeros_like(self.Rs) if rv is None else np.asarray(rv, dtype=np.float64)
        n = len(self.P)
        for arr, name in ((self.Bo, "Bo"), (self.Bw, "Bw"), (self.Bg, "Bg"), (self.mu_o, "mu_o_cP"),
                          (self.mu_w, "mu_w_cP"), (self.mu_g, "mu_g_cP"), (self.Rs, "Rs_m3m3")):
            if len(arr) != n:
                raise ValueError(f"PVT: column lengths differ (pressure vs {name})")
        if len(self.Rv) != n:
            raise ValueError("PVT: column lengths differ (pressure vs Rv_m3m3)")
        if n < 2 or np.any(np.diff(self.P) <= 0):
            raise ValueError("PVT: pressure_MPa must hold >= 2 strictly increasing values")

    def _interp(self, x, xp, fp):
        return np.interp(x, xp, fp, left=fp[0], right=fp[-1])

    def eval(self, P_MPa: np.ndarray) -> dict:
        p = np.asarray(P_MPa).astype(np.float64)
        return {"Bo": self._interp(p, self.P, self.Bo), "Bw": self._interp(p, self.P, self.Bw),
                "Bg": self._interp(p, self.P, self.Bg), "mu_o": self._interp(p, self.P, self.mu_o),
                "mu_w": self._interp(p, self.P, self.mu_w), "mu_g": self._interp(p, self.P, self.mu_g),
                "Rs": self._interp(p, self.P, self.Rs), "Rv": self._interp(p, self.P, self.Rv)}

    def device_table(self) -> np.ndarray:
        """9 x n array (pressure then Bo Bw Bg mu_o mu_w mu_g Rs Rv) for the HIP library."""
        return np.stack([self.P, self.Bo, self.Bw, self.Bg, self.mu_o, self.mu_w, self.mu_g, self.Rs, self.Rv])
