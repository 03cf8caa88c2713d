"""Host-side mirror of the reference's DiagnosticsList models (DiagnosticsList.C:111-151),
computed from the device reductions of pic_diag_reduce and pic_cell_moments.

Each function returns the numbers the reference prints in its info() line."""
from __future__ import annotations

import numpy as np

from . import picmi

KB = 1.380649e-23
QE = 1.602176634e-19
EPS0 = 8.8541878128e-12


def temperature(cloud: "picmi.Cloud", account_for_drift: bool = True):
    """TemperatureInfo::info (TemperatureInfo.C:86-138): (system temperature [K], per-species
    temperatures [K]) from sum w U^2, sum w U and sum w per species."""
    d = cloud.diag()
    ns = len(cloud.species)
    masses = [s["mass"] for s in cloud.species]
    n_tot = sum(d.sum_w[s] for s in range(ns))
    if n_tot <= 0:
        return 0.0, [0.0] * ns
    t_sys = 0.0
    per = []
    for s in range(ns):
        n = d.sum_w[s]
        if n <= 0:
            per.append(0.0)
            continue
        vd = np.array(d.sum_wU[s][:])
        vd2 = float(vd @ vd) if account_for_drift else 0.0
        t_sys += masses[s] * (d.sum_wU2[s] - vd2 / n)
        per.append(masses[s] / (3.0 * KB) * (d.sum_wU2[s] / n - vd2 / (n * n)))
    t_sys /= 3.0 * KB * n_tot
    return t_sys, per


def kinetic_energy(cloud: "picmi.Cloud", relativistic: bool = False):
    """KineticEnergyInfo (KineticEnergyInfo.C:69-82): per-species kinetic energy [J].  The
    relativistic branch sums (gamma - 1) m c^2 without the parcel weight, as the reference."""
    d = cloud.diag()
    ns = len(cloud.species)
    if relativistic:
        return [d.kinetic_energy_relativistic[s] for s in range(ns)]
    return [d.kinetic_energy[s] for s in range(ns)]


def field_energy(cloud: "picmi.Cloud", from_parcel_distribution: bool = False):
    """FieldEnergyInfo::info (FieldEnergyInfo.C:80-131): total field energy [J] and the
    per-species parcel terms (from_parcel_distribution: 0.5 sum q w phi(cell) per species
    plus the space-charge term sum phi rho_space V)."""
    if not from_parcel_distribution:
        return cloud.diag().field_energy, []
    phi = cloud.get_field(picmi.FIELD_PHI)
    space = cloud.get_field(picmi.FIELD_SPACE_CHARGE)
    _, V = cloud.geometry()
    per = []
    for s in range(len(cloud.species)):
        m = cloud.cell_moments(s)
        per.append(float(np.sum(m[:, 10] * V * phi)))
    total = 0.5 * sum(per) + float(np.sum(phi * space * V))
    return total, per


def plasma_frequency(cloud: "picmi.Cloud", electron_species: int, total_volume: float | None = None):
    """PlasmaFrequencyInfo::info (PlasmaFrequencyInfo.C:84-103): sqrt(n q^2 / (eps0 m)) from
    the electron weight over the domain volume (not a measured oscillation)."""
    d = cloud.diag()
    if total_volume is None:
        total_volume = float(np.sum(cloud.geometry()[1]))
    sp = cloud.species[electron_species]
    n = d.sum_w[electron_species] / total_volume
    return float(np.sqrt(n * sp["charge"] ** 2 / EPS0 / sp["mass"]))
