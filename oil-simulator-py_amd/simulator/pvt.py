"""PVT table of the black-oil model: the reference's `PVTTable` interface
(src/simulator/pvt.py:18-64): the JSON schema, the attribute names P, Bo, Bw,
Bg, mu_o, mu_w, mu_g, Rs, Rv, `eval` (linear interpolation in pressure, held at
the end values outside the table) and ValueError on an inconsistent table.

The table is one (9, n) FP64 array whose rows follow `COLUMNS`; that array is
also what the HIP library receives (impes_model_create), where the IMPES step
evaluates it on the device. `eval` is the host-side lookup for callers that
hold numpy arrays.
"""
from __future__ import annotations

import json

import numpy as np

# row of the table -> (JSON key, attribute, eval() key, required); row 0 is the
# pressure axis in MPa. Rv (vaporised oil) may be absent: a column of zeros.
COLUMNS = (
    ("pressure_MPa", "P", None, True),
    ("Bo", "Bo", "Bo", True),
    ("Bw", "Bw", "Bw", True),
    ("Bg", "Bg", "Bg", True),
    ("mu_o_cP", "mu_o", "mu_o", True),
    ("mu_w_cP", "mu_w", "mu_w", True),
    ("mu_g_cP", "mu_g", "mu_g", True),
    ("Rs_m3m3", "Rs", "Rs", True),
    ("Rv_m3m3", "Rv", "Rv", False),
)


def _load(json_path: str) -> dict:
    with open(json_path, "r", encoding="utf-8") as f:
        return json.load(f)


class PVTTable:
    def __init__(self, json_path: str | None = None, data: dict | None = None):
        src = _load(json_path) if data is None else data
        axis = np.asarray(src["pressure_MPa"], dtype=np.float64).ravel()
        if axis.size < 2 or not np.all(axis[1:] > axis[:-1]):
            raise ValueError("PVT: pressure_MPa must hold >= 2 strictly increasing values")
        table = np.empty((len(COLUMNS), axis.size))
        for row, (key, _, _, required) in enumerate(COLUMNS):
            raw = src.get(key)
            if raw is None and not required:
                table[row] = 0.0
                continue
            col = np.asarray(raw if raw is not None else [], dtype=np.float64).ravel()
            if col.size != axis.size:
                raise ValueError(f"PVT: column {key} holds {col.size} values, pressure_MPa {axis.size}")
            table[row] = col
        self._table = table

    def eval(self, P_MPa) -> dict:
        """Every property at the pressures P_MPa (np.interp, clamped to the end values)."""
        p = np.asarray(P_MPa, dtype=np.float64)
        axis = self._table[0]
        return {name: np.interp(p, axis, self._table[row])
                for row, (_, _, name, _) in enumerate(COLUMNS) if name is not None}

    def device_table(self) -> np.ndarray:
        """The (9, n) table in COLUMNS order, as the HIP library takes it."""
        return self._table.copy()


def _column(row: int):
    return property(lambda self: self._table[row], doc=f"table column {COLUMNS[row][0]}")


for _row, (_key, _attr, _name, _req) in enumerate(COLUMNS):
    setattr(PVTTable, _attr, _column(_row))
del _row, _key, _attr, _name, _req
