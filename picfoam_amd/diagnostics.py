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


def velocity(cloud: "picmi.Cloud"):
    """VelocityInfo::info (VelocityInfo.C:58-80): the drift velocity sum w U / sum w of each
    species with parcels (None for the others)."""
    d = cloud.diag()
    return [np.array(d.sum_wU[s][:]) / d.sum_w[s] if d.sum_w[s] > 0 else None for s in range(len(cloud.species))]


def momentum(cloud: "picmi.Cloud", relativistic: bool = False):
    """MomentumInfo::info (MomentumInfo.C:58-106): (total momentum, per-species momentum)
    [N s], sum m w U, or sum m w gamma U in the relativistic branch."""
    d = cloud.diag()
    per = []
    for s, sp in enumerate(cloud.species):
        if relativistic:
            per.append(np.array(d.momentum_relativistic[s][:]))
        else:
            per.append(sp["mass"] * np.array(d.sum_wU[s][:]))
    return np.sum(per, axis=0), per


def cloud_composition(cloud: "picmi.Cloud", total_volume: float | None = None):
    """CloudCompositionInfo::info (CloudCompositionInfo.C:66-100): the number of real
    particles, and per species (parcels, particles, number density [m^-3])."""
    d = cloud.diag()
    if total_volume is None:
        total_volume = float(np.sum(cloud.geometry()[1]))
    per = [(int(d.n_parcels[s]), d.sum_w[s], d.sum_w[s] / total_volume) for s in range(len(cloud.species))]
    return sum(p[1] for p in per), per


def charge(cloud: "picmi.Cloud"):
    """ChargeInfo::info (ChargeInfo.C:60-72): sum over cells of (rhoCharge + spaceCharge) /
    V, as the reference writes it (a charge density divided by the cell volume, not times
    it)."""
    rho = cloud.get_field(picmi.FIELD_RHO)
    space = cloud.get_field(picmi.FIELD_SPACE_CHARGE)
    V = cloud.geometry()[1]
    return float(np.sum((rho + space) / V))


def cell_occupancy(cloud: "picmi.Cloud"):
    """CellOccupancyInfo::info (CellOccupancyInfo.C:60-95): (average, maximum, minimum)
    parcels per cell over every species (picRhoN)."""
    n = cloud.cell_moments(-1)[:, 2]
    return float(np.sum(n)) / len(n), float(np.max(n)), float(np.min(n))


def isotropization(cloud: "picmi.Cloud", account_for_drift: bool = True):
    """IsotropizationInfo::info (IsotropizationInfo.C:68-118): per species with parcels
    (T_x, T_yz, T_yz / T_x) [K], T_x = m/k (<w u_x^2>/<w> - drift_x^2), T_yz = m/(2k)
    (<w (u_y^2 + u_z^2)>/<w> - drift_y^2 - drift_z^2)."""
    d = cloud.diag()
    out = []
    for s, sp in enumerate(cloud.species):
        n = d.sum_w[s]
        if n <= 0:
            out.append(None)
            continue
        vd = np.array(d.sum_wU[s][:]) / n
        d2x = vd[0] ** 2 if account_for_drift else 0.0
        d2yz = vd[1] ** 2 + vd[2] ** 2 if account_for_drift else 0.0
        tx = sp["mass"] / KB * (d.sum_wUU[s][0] / n - d2x)
        tyz = sp["mass"] / (2.0 * KB) * ((d.sum_wUU[s][1] + d.sum_wUU[s][2]) / n - d2yz)
        out.append((tx, tyz, tyz / tx))
    return out


def unidirectional_velocity(cloud: "picmi.Cloud", direction_mask=(1.0, 0.0, 0.0)):
    """UnidirectionalVelocityInfo::info (UnidirectionalVelocityInfo.C:66-100): per species
    with parcels (sum w U*U . mask) / sum w, the mean square velocity along the masked
    directions."""
    d = cloud.diag()
    m = np.asarray(direction_mask, np.float64)
    return [float(np.dot(d.sum_wUU[s][:], m)) / d.sum_w[s] if d.sum_w[s] > 0 else None
            for s in range(len(cloud.species))]


def debye_length(cloud: "picmi.Cloud", account_for_drift: bool = True, total_volume: float | None = None):
    """DebyeLengthInfo::info (DebyeLengthInfo.C:101-146): sqrt(1 / sum n q^2 / (eps0 k T))
    over the charged species with parcels, T from each species' sums (the drift removed when
    accountForDrift), n = sum w / volume, q = sum q w / sum w; 0 when no species
    contributes.  Returns (Debye length [m], mean cell length scale [m]): the reference
    warns when the second exceeds the first."""
    d = cloud.diag()
    if total_volume is None:
        total_volume = float(np.sum(cloud.geometry()[1]))
    if sum(d.sum_w[s] for s in range(len(cloud.species))) <= 0:
        return 0.0, None
    inv = 0.0
    for s, sp in enumerate(cloud.species):
        n_w = d.sum_w[s]
        if sp["charge"] == 0.0 or n_w <= 0.0:
            continue
        vd = np.array(d.sum_wU[s][:])
        vd2 = float(vd @ vd) if account_for_drift else 0.0
        T = sp["mass"] * (d.sum_wU2[s] - vd2 / n_w) / (3.0 * KB * n_w)
        n = n_w / total_volume
        q = d.charge[s] / n_w
        inv += n * q * q / (EPS0 * KB * T)
    lam = float(np.sqrt(1.0 / inv)) if inv > 0.0 else 0.0
    V = cloud.geometry()[1]
    return lam, float(np.mean(np.cbrt(V)))


def ionization_states(cloud: "picmi.Cloud", limits: dict):
    """IonizationInfo::info (IonizationInfo.C:60-105): per ion species the parcel count of
    each charge state up to its ionizationLimit.  Every parcel of this build carries its
    species' charge (charge state 1), so the counts are the parcel numbers in state 1."""
    d = cloud.diag()
    out = {}
    for s, lim in limits.items():
        states = [0] * int(lim)
        if lim >= 1:
            states[0] = int(d.n_parcels[s])
        out[s] = states
    return out


def print_parcel_sums(cloud: "picmi.Cloud"):
    """PrintParcelInfo with `sums` (PrintParcelInfo.C:80-110): the sums of every parcel's
    position and velocity."""
    sp, su = np.zeros(3), np.zeros(3)
    for s in range(len(cloud.species)):
        if cloud.count(s) == 0:
            continue
        x = cloud.positions(s)
        _, t, U, _ = cloud.get_particles(s)
        live = t >= 0
        sp += x[live].sum(0)
        su += U[live].sum(0)
    return sp, su


def dump_parcels(cloud: "picmi.Cloud", path: str):
    """DumpInfo (DumpInfo.C:63-66) -> PICCloud::dumpParticle (PICCloud.C:1523-1549): one line
    per parcel, "x X Y Z u Ux Uy Uy t typeId n nParticle c chargeModifier".  The reference
    writes U.y twice (PICCloud.C:1540-1541); the file keeps that layout."""
    with open(path, "w") as f:
        for s, sp in enumerate(cloud.species):
            if cloud.count(s) == 0:
                continue
            x = cloud.positions(s)
            _, t, U, _ = cloud.get_particles(s)
            w = sp["nEquivalentParticles"]
            for i in np.nonzero(t >= 0)[0]:
                f.write(f"x {x[i, 0]:.17g} {x[i, 1]:.17g} {x[i, 2]:.17g} u {U[i, 0]:.17g} {U[i, 1]:.17g} "
                        f"{U[i, 1]:.17g} t {s} n {w:.17g} c 1\n")


# DiagnosticInfo models by their reference names
MODELS = ("Temperature", "KineticEnergy", "FieldEnergy", "PlasmaFrequency", "Velocity", "Momentum",
          "CloudComposition", "Charge", "CellOccupancy", "Isotropization", "UnidirectionalVelocity", "DebyeLength",
          "Ionization", "PrintParcel", "Dump")


class DiagnosticsList:
    """DiagnosticsList (DiagnosticsList.C:40-151): the models of Diagnostics.modelList, each
    with its <model>Coeffs.  A model runs when its diagnosticControl says so
    (DiagnosticInfo.C:90-94): timeStep (the default) every diagnosticInterval steps
    (every step when the interval is 0 or 1), or runTime every diagnosticInterval seconds
    of simulated time."""

    def __init__(self, cloud: "picmi.Cloud", model_list, coeffs: dict | None = None, dt: float = 0.0,
                 electron_species: int | None = None):
        unknown = [m for m in model_list if m not in MODELS]
        if unknown:
            raise picmi.PicError(-2, f"Unknown DiagnosticInfo type(s) {unknown}. Valid types: {list(MODELS)}")
        self.cloud = cloud
        self.models = list(model_list)
        self.coeffs = {m: dict((coeffs or {}).get(m + "Coeffs", {})) for m in self.models}
        self.dt = dt
        self.electron_species = electron_species
        self._last = {m: None for m in self.models}

    def _due(self, m, step_index, time):
        c = self.coeffs[m]
        ctl = c.get("diagnosticControl", "timeStep")
        iv = float(c.get("diagnosticInterval", 0))
        if ctl == "timeStep":
            return iv <= 1 or step_index % int(iv) == 0
        if ctl in ("runTime", "adjustableRunTime"):
            k = int(np.floor((time + 0.5 * self.dt) / iv))
            if self._last[m] is None or k > self._last[m]:
                self._last[m] = k
                return True
            return False
        raise picmi.PicError(-1, f"diagnosticControl {ctl} is not supported (timeStep, runTime)")

    def info(self, step_index: int, time: float = 0.0, dump_dir: str | None = None, time_name: str = ""):
        """The values of the models due at this step, by model name."""
        out = {}
        cl = self.cloud
        for m in self.models:
            if not self._due(m, step_index, time):
                continue
            c = self.coeffs[m]
            drift = bool(c.get("accountForDrift", True))
            if m == "Temperature":
                out[m] = temperature(cl, drift)
            elif m == "KineticEnergy":
                out[m] = kinetic_energy(cl, bool(c.get("relativistic", False)))
            elif m == "FieldEnergy":
                out[m] = field_energy(cl, bool(c.get("fromDistribution", False)))
            elif m == "PlasmaFrequency":
                out[m] = plasma_frequency(cl, self.electron_species)
            elif m == "Velocity":
                out[m] = velocity(cl)
            elif m == "Momentum":
                out[m] = momentum(cl, bool(c.get("relativistic", False)))
            elif m == "CloudComposition":
                out[m] = cloud_composition(cl)
            elif m == "Charge":
                out[m] = charge(cl)
            elif m == "CellOccupancy":
                out[m] = cell_occupancy(cl)
            elif m == "Isotropization":
                out[m] = isotropization(cl, drift)
            elif m == "UnidirectionalVelocity":
                out[m] = unidirectional_velocity(cl, c["direction"])
            elif m == "DebyeLength":
                out[m] = debye_length(cl, drift)
            elif m == "Ionization":
                # a sub-dictionary per ion species (by index here) with its ionizationLimit
                out[m] = ionization_states(cl, {int(k): v["ionizationLimit"] for k, v in c.items()
                                                if isinstance(v, dict)})
            elif m == "PrintParcel":
                out[m] = print_parcel_sums(cl)
            elif m == "Dump":
                import os
                path = os.path.join(dump_dir or ".", f"parcels_pic_{time_name or step_index}.obj")
                dump_parcels(cl, path)
                out[m] = path
        return out
