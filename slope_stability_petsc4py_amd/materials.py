"""Material rows and their expansion to quadrature points (mesh/materials.py:11-101).

``MaterialSpec`` holds one Mohr-Coulomb material (c0 [kPa], phi and psi [deg], Young [kPa],
Poisson, unit weights [kN/m^3]); the elastic moduli follow the reference's formulas.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class MaterialSpec:
    c0: float
    phi: float
    psi: float
    young: float
    poisson: float
    gamma_sat: float
    gamma_unsat: float

    @property
    def shear(self) -> float:
        return self.young / (2.0 * (1.0 + self.poisson))

    @property
    def bulk(self) -> float:
        return self.young / (3.0 * (1.0 - 2.0 * self.poisson))

    @property
    def lame(self) -> float:
        return self.bulk - 2.0 * self.shear / 3.0


def as_specs(materials) -> list[MaterialSpec]:
    if isinstance(materials, (dict, MaterialSpec)):
        materials = [materials]
    out = []
    for m in materials:
        if isinstance(m, MaterialSpec):
            out.append(m)
        else:
            keys = ("c0", "phi", "psi", "young", "poisson", "gamma_sat", "gamma_unsat")
            out.append(MaterialSpec(**{k: float(m[k]) for k in keys}))
    return out


def expand_materials(material_id: np.ndarray, n_q: int, materials, saturation: np.ndarray | None = None):
    """Per-point (c0, phi[rad], psi[rad], shear, bulk, lame, gamma) arrays, element-major points."""
    specs = as_specs(materials)
    mid = np.asarray(material_id, dtype=np.int64).ravel()
    if len(specs) == 1 and mid.size and mid.max() > 0:
        specs = specs * (int(mid.max()) + 1)
    if mid.size and mid.max() >= len(specs):
        raise IndexError(f"Material identifier {int(mid.max())} requires at least {int(mid.max()) + 1} material rows")
    tab = np.array([[s.c0, np.deg2rad(s.phi), np.deg2rad(s.psi), s.shear, s.bulk, s.lame, s.gamma_sat,
                     s.gamma_unsat] for s in specs])
    per_e = tab[mid]
    per_q = np.repeat(per_e, n_q, axis=0)
    sat = np.ones(per_q.shape[0], dtype=bool) if saturation is None else np.asarray(saturation, dtype=bool).ravel()
    gamma = np.where(sat, per_q[:, 6], per_q[:, 7])
    return (per_q[:, 0].copy(), per_q[:, 1].copy(), per_q[:, 2].copy(), per_q[:, 3].copy(), per_q[:, 4].copy(),
            per_q[:, 5].copy(), gamma)
