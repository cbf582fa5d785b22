"""Cartesian reservoir grid (reference: src/simulator/reservoir.py:8-80).

Fields are float64 tensors on the ROCm device, shape (nx, ny, nz) with z
fastest in memory. Heterogeneous fields are set by overwriting
permeability_x/y/z (m^2), porosity and porous_volume before the Simulator
is constructed, exactly as with the reference.
"""
from __future__ import annotations

import torch

MD_TO_M2 = 9.869233e-16


class Reservoir:
    def __init__(self, config, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.dimensions = list(config.get("dimensions", [10, 10, 1]))
        self.nx, self.ny, self.nz = (int(v) for v in self.dimensions)
        self.grid_size = tuple(float(v) for v in config.get("grid_size", [10.0, 10.0, 10.0]))
        self.dx, self.dy, self.dz = self.grid_size
        porosity_value = config.get("porosity", 0.2)
        f64 = dict(dtype=torch.float64, device=self.device)
        self.porosity = torch.full(self.dimensions, float(porosity_value), **f64)
        k_h = float(config.get("permeability", 100.0))
        k_v = k_h * float(config.get("k_vertical_fraction", 0.1))
        self.permeability_x = torch.full(self.dimensions, k_h * MD_TO_M2, **f64)
        self.permeability_y = torch.full(self.dimensions, k_h * MD_TO_M2, **f64)
        self.permeability_z = torch.full(self.dimensions, k_v * MD_TO_M2, **f64)
        self.rock_compressibility = float(config.get("c_rock", 1e-5)) / 1e6
        self.grid_size = torch.tensor(self.grid_size, dtype=torch.float64, device=self.device)
        self.cell_volume = self.dx * self.dy * self.dz
        self.porous_volume = torch.full(self.dimensions, self.cell_volume * float(porosity_value), **f64)

    @property
    def permeability_tensors(self):
        return self.permeability_x, self.permeability_y, self.permeability_z

    def get_cell_indices(self, i, j, k):
        """reservoir.py:70-80 (x fastest; the solver itself uses z fastest)."""
        return i + j * self.nx + k * self.nx * self.ny

    def set_fields(self, perm_x_md=None, perm_y_md=None, perm_z_md=None, porosity=None):
        """Overwrite fields from arrays in mD / fraction and refresh the pore volume."""
        f64 = dict(dtype=torch.float64, device=self.device)
        for name, v in (("permeability_x", perm_x_md), ("permeability_y", perm_y_md),
                        ("permeability_z", perm_z_md)):
            if v is not None:
                setattr(self, name, torch.as_tensor(v, **f64).reshape(self.dimensions) * MD_TO_M2)
        if porosity is not None:
            self.porosity = torch.as_tensor(porosity, **f64).reshape(self.dimensions)
            self.porous_volume = self.porosity * self.cell_volume
