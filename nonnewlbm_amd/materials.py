"""Materials and the physical -> lattice conversions of sim_tcpc.

``materialByName`` (solver/sim_tcpc/sim_tcpc.cu:52-68) knows newtonian, sx and
gx; the benchmark configurations add a blood-like Newtonian fluid and two
Carreau-Yasuda sets.  ``lattice_units`` restates sim_tcpc.cu:795-814:

    dt        = nu_lbm / nu_inf * dl^2
    nu0_lbm   = mu0 / rho / dl^2 * dt       (CY only, else nu_lbm)
    lambda    = lambda_phys / dt            (CY only, else 1)
    flux_lbm  = Q[ml/s] * 1e-6 / dl^3 * dt
"""
from __future__ import annotations

import dataclasses


@dataclasses.dataclass(frozen=True)
class Material:
    name: str
    density: float        # kg/m^3
    nu_inf: float         # m^2/s
    mu0: float            # Pa s
    lam: float            # s
    a: float
    n: float
    carreau_yasuda: bool


MATERIALS = {
    # sim_tcpc.cu:54-66
    "newtonian": Material("newtonian", 1000.0, 1.0e-6, 1.0e-3, 1.0, 2.0, 1.0, False),
    "sx": Material("sx", 1151.0, 3.72863999716223e-3 / 1151.0, 3.32933884442747e-2, 1.20785676226298,
                   0.925354152818718, 0.463126647, True),
    "gx": Material("gx", 1088.0, 4.85563692580026e-3 / 1088.0, 1.58517801201338e-2, 0.0280506051570583,
                   2.34147778848562, 0.463126647, True),
    # the benchmark materials (density 1060 kg/m^3)
    "blood-newtonian": Material("blood-newtonian", 1060.0, 0.0035 / 1060.0, 0.0035, 1.0, 2.0, 1.0, False),
    "cy-a": Material("cy-a", 1060.0, 0.00345 / 1060.0, 0.056, 3.313, 2.0, 0.3568, True),
    "cy-b": Material("cy-b", 1060.0, 0.0035 / 1060.0, 0.16, 8.2, 0.64, 0.2128, True),
}


def material_by_name(name: str) -> Material:
    try:
        return MATERIALS[name]
    except KeyError:
        raise ValueError(f"Unknown material '{name}'. Use {', '.join(MATERIALS)}.") from None


@dataclasses.dataclass(frozen=True)
class LatticeUnits:
    dt: float
    nu_lbm: float
    nu0: float
    lam: float
    a: float
    n: float
    cy: bool

    def flux(self, mls: float, dl: float) -> float:
        return mls * 1e-6 / (dl * dl * dl) * self.dt


def lattice_units(material: Material, dl: float, lbm_viscosity: float) -> LatticeUnits:
    if not dl > 0.0:
        raise ValueError("Geometry spacing must be positive.")
    dt = lbm_viscosity / material.nu_inf * dl * dl
    if material.carreau_yasuda:
        nu0 = material.mu0 / material.density / dl / dl * dt
        return LatticeUnits(dt, lbm_viscosity, nu0, material.lam / dt, material.a, material.n, True)
    return LatticeUnits(dt, lbm_viscosity, lbm_viscosity, 1.0, 2.0, 1.0, False)


def case_kwargs(material: Material, dl: float, lbm_viscosity: float, cy_model: int | None = None) -> dict:
    """CaseDesc fields of a material (dt, viscosities, CY parameters)."""
    from . import _abi
    u = lattice_units(material, dl, lbm_viscosity)
    kw = dict(dl=dl, dt=u.dt, nu_ref=lbm_viscosity, cy_enabled=1 if u.cy else 0,
              cy_model=(_abi.CY_NEQ if cy_model is None else cy_model) if u.cy else _abi.CY_OFF,
              cy_nu0=u.nu0, cy_lambda=u.lam, cy_a=u.a, cy_n=u.n)
    return kw
