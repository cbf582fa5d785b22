"""Wells (reference: src/simulator/well.py:10-132)."""
from __future__ import annotations

import math

SURFACE_PHASES = {"oil": 0, "water": 1, "gas": 2, "liquid": 3}
LIMIT_KEYS = ("wopr", "wlpr", "wgpr", "liqr")


class Well:
    def __init__(self, name, well_type, i, j, k, radius, control_type, control_value, reservoir_dimensions,
                 injected_phase: str = "water", rate_type: str = "reservoir", surface_phase: str | None = None,
                 limits: dict | None = None, bhp_min: float | None = None):
        self.name = name
        self.type = well_type
        self.i, self.j, self.k = int(i), int(j), int(k)
        self.radius = radius
        self.control_type = control_type
        self.control_value = control_value
        self.injected_phase = injected_phase
        self.rate_type = rate_type
        self.surface_phase = surface_phase
        self.limits = limits or {}
        self.bhp_min = bhp_min
        nx, ny, nz = reservoir_dimensions
        self.cell_index = (self.i, self.j, self.k)
        self.cell_index_flat = self.i + self.j * nx + self.k * nx * ny
        # Peaceman index with the reference's assumed 100 mD and dx = dy = 1 (well.py:44-51)
        dx = dy = 1.0
        ro = 0.28 * ((dx ** 2 + dy ** 2) ** 0.5) / 2
        self.well_index = 2 * math.pi * 100.0 / (math.log(ro / radius) + 1e-12)
        self.last_q_total = None
        self.last_surface_rates = {}
        self.last_mode = None

    def __str__(self):
        return (f"Well(name={self.name}, type={self.type}, pos=({self.i},{self.j},{self.k}), "
                f"control={self.control_type}:{self.control_value})")


class WellManager:
    def __init__(self, well_configs, reservoir):
        self.wells = []
        for w in well_configs:
            if "coordinates" in w:
                i, j, k = w["coordinates"]
            elif "i" in w and "j" in w and "k" in w:
                i, j, k = w["i"], w["j"], w["k"]
            else:
                raise ValueError(f"invalid well coordinates: {w}")
            if "control" in w:
                ct, cv = w["control"]["type"], w["control"]["value"]
            elif "control_type" in w and "control_value" in w:
                ct, cv = w["control_type"], w["control_value"]
            else:
                raise ValueError(f"invalid well control: {w}")
            self.wells.append(Well(name=w["name"], well_type=w["type"], i=i, j=j, k=k, radius=w["radius"],
                                   control_type=ct, control_value=cv, reservoir_dimensions=reservoir.dimensions,
                                   injected_phase=w.get("injected_phase", "water"),
                                   rate_type=w.get("rate_type", "reservoir"), surface_phase=w.get("surface_phase"),
                                   limits=w.get("limits"), bhp_min=w.get("bhp_min")))

    def add_well(self, well):
        self.wells.append(well)

    def get_wells(self):
        return self.wells

    def get_well_indices_flat(self):
        import torch

        return torch.tensor([w.cell_index_flat for w in self.wells], dtype=torch.long)
