"""Synthetic plasmas of the benchmark configurations (meshes, species, models).

Each builder returns ``(mesh, species, models, extra)`` ready for ``picmi.Cloud``; the
sizes can be scaled down for tests.  Choices the configuration text leaves open (box
sizes, time steps, boundary conditions) are fixed here and documented with the value.
"""
from __future__ import annotations

import numpy as np

from . import meshgen

QE = 1.602176634e-19
ME = 9.1093837015e-31
EPS0 = 8.8541878128e-12
KB = 1.380649e-23
EV_K = QE / KB  # kelvin per eV
AMU = 1.66053906660e-27


def plasma_frequency(n, q=QE, m=ME):
    return np.sqrt(n * q * q / (EPS0 * m))


def cold_plasma_oscillation(ncell=256, ppc=100, L=0.1, n=1e14, amplitude=0.01, steps_per_wp=0.05,
                            weighting="CellAverage", precision="fp64"):
    """Periodic 1-D slab of ``ncell`` cubic hex cells along z, electrons uniformly spaced
    and displaced by amplitude*L/(2 pi) * sin(2 pi z/L) (a 1 % density perturbation),
    cold, over an immobile neutralising background (space charge)."""
    dz = L / ncell
    mesh = meshgen.hex_box(1, 1, ncell, dz, dz, L, cyclic=(True, True, True))
    N = ncell * ppc
    vol = dz * dz * L
    w = n * vol / N
    wp = plasma_frequency(n)
    dt = steps_per_wp / wp
    species = [{"name": "e", "mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": N}]
    models = {"ParticlePusher": "Boris", "FieldWeigthing": weighting, "ChargeDistribution": weighting,
              "MaxwellSolver": "ElectroStatic", "precision": precision, "poissonTolerance": 1e-10}
    z0 = (np.arange(N) + 0.5) * L / N
    k = 2 * np.pi / L
    # displacement xi = A sin(kz) gives delta n/n = -A k cos(kz); A k = amplitude
    z = z0 + amplitude / k * np.sin(k * z0)
    z = np.mod(z, L)
    x = np.stack([np.full(N, 0.5 * dz), np.full(N, 0.5 * dz), z], axis=1)
    cell = np.minimum((z / dz).astype(np.int64), ncell - 1).astype(np.int32)
    extra = {"x": x, "cell": cell, "U": np.zeros((N, 3)), "ids": np.arange(N, dtype=np.uint32), "dt": dt, "wp": wp,
             "space_charge": np.full(ncell, n * QE), "k": k, "L": L, "dz": dz, "z0": z0}
    return mesh, species, models, extra


def thermalisation(ncell=32, ppc=64, n=1e18, Te_eV=10.0, Ti_eV=1.0, mass_ratio=1836.0, coulomb_log=10.0,
                   box=None, dt=None, precision="fp64"):
    """Two-species thermalisation: all-cyclic ncell^3 hex box, ppc electrons and ppc
    singly charged ions (m_i = mass_ratio m_e) per cell, Maxwellian at Te / Ti, Boris
    push, Nanbu + Takizuka-Abe Coulomb collisions with a fixed Coulomb logarithm and no
    field solve (MaxwellSolver none).

    dt defaults to 0.2 / (5000 nu_eq): 5000 steps relax the temperature difference by
    exp(-0.4).  Nanbu's scattering needs this resolution: at 2 / (5000 nu_eq) the e-e and
    e-i scattering parameter s reaches O(1) for slow electrons and the relaxation runs
    about 40 % slower than Spitzer (checked on the oracle).  With no field solve the plasma is homogeneous and the box size only sets how far a
    parcel moves per step; it defaults to cells of side v_th,e dt / 0.3, i.e. a thermal
    electron crosses 0.3 cell per step (the reference's warnCellTrajectory 0.3)."""
    mi = mass_ratio * ME
    nu = spitzer_nu(n, Te_eV, Ti_eV, ME, mi, coulomb_log)
    if dt is None:
        dt = 0.2 / nu / 5000
    vth_e = np.sqrt(Te_eV * QE / ME)
    if box is None:
        box = ncell * vth_e * dt / 0.3
    h = box / ncell
    mesh = meshgen.hex_box(ncell, ncell, ncell, box, box, box, cyclic=(True, True, True))
    vol_cell = h ** 3
    w = n * vol_cell / ppc
    N = ncell ** 3 * ppc
    species = [{"name": "e", "mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": N},
               {"name": "i", "mass": mi, "charge": QE, "nEquivalentParticles": w, "capacity": N}]
    models = {"ParticlePusher": "Boris", "MaxwellSolver": "none", "CoulombCollisionModel": "Nanbu",
              "PairingAlgorithm": "TakizukaAbe", "coulombLog": {(0, 0): coulomb_log, (0, 1): coulomb_log,
                                                                 (1, 1): coulomb_log},
              "precision": precision}
    extra = {"Te_K": Te_eV * EV_K, "Ti_K": Ti_eV * EV_K, "dt": dt, "nu_eq": nu, "ppc": ppc, "box": box}
    return mesh, species, models, extra


def binary_mixture(ncell=100, ppc_ne=26, ppc_ar=13, n=1e22, T_ne=1000.0, T_ar=500.0, h=1.875e-6, dt=1e-10,
                   cross_section="HardSphere", precision="fp64"):
    """Neon and argon relaxing by hard-sphere (or variable-hard-sphere) collisions in an
    all-cyclic ncell^3 hex box of cells of side h (the reference's
    thermalization_binaryCollision gas: 1e22 m^-3 of each at 1000 K and 500 K, diameters
    3.08e-10 and 3.76e-10 m), with parcel weights set by the parcels per cell;
    IsotropicScattering with the Nanbu weight correction, no field."""
    mesh = meshgen.hex_box(ncell, ncell, ncell, ncell * h, ncell * h, ncell * h, cyclic=(True, True, True))
    vol = h ** 3
    N = ncell ** 3
    species = [{"name": "Ne", "mass": 3.375350746e-26, "charge": 0.0, "nEquivalentParticles": n * vol / ppc_ne,
                "capacity": N * ppc_ne, "diameter": 3.08e-10, "omega": 1.0},
               {"name": "Ar", "mass": 6.681789475e-26, "charge": 0.0, "nEquivalentParticles": n * vol / ppc_ar,
                "capacity": N * ppc_ar, "diameter": 3.76e-10, "omega": 1.0}]
    models = {"ParticlePusher": "Boris", "MaxwellSolver": "none", "BinaryCollisionModel": "IsotropicScattering",
              "BinaryCollisionCoeffs": {"TotalCrossSectionModel": cross_section, "Tref": 273.0,
                                        "WeightCorrection": "Nanbu" if ppc_ne != ppc_ar else "none"},
              "precision": precision}
    extra = {"T": (T_ne, T_ar), "n": (n, n), "dt": dt, "ppc": (ppc_ne, ppc_ar)}
    return mesh, species, models, extra


def spitzer_nu(n, Te_eV, Ti_eV, me, mi, lnL):
    """nu_eq of dT_e/dt = nu_eq (T_i - T_e) for Z=1, n_e = n_i (NRL formulary form)."""
    Te = Te_eV * QE; Ti = Ti_eV * QE
    return (2.0 * n * QE ** 4 * lnL / (3.0 * (2 * np.pi) ** 1.5 * EPS0 ** 2 * me * mi *
                                       (Te / me + Ti / mi) ** 1.5))


def magnetised_tets(n_lattice=50, ppc=None, n_total=50_000_000, T_eV=5.0, Bz=0.1, Ex=1e4, L=0.01, jitter=0.2,
                    precision="mixed", pusher="HigueraCary"):
    """Magnetised electron push on a jittered tet lattice of the 1 cm cube (6 n^3 cells),
    uniform Bz and Ex (MaxwellSolver none), Maxwellian electrons."""
    mesh = meshgen.tet_lattice(n_lattice, L, jitter=jitter, seed=2024)
    ncell = mesh["n_cells"]
    if ppc is None:
        ppc = max(1, int(round(n_total / ncell)))
    N = ncell * ppc
    wc = QE * Bz / ME
    dt = 0.2 / wc
    species = [{"name": "e", "mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": N}]
    models = {"ParticlePusher": pusher, "MaxwellSolver": "none", "precision": precision}
    extra = {"E": np.tile([Ex, 0.0, 0.0], (ncell, 1)), "B": np.tile([0.0, 0.0, Bz], (ncell, 1)), "dt": dt,
             "T_K": T_eV * EV_K, "ppc": ppc, "vExB": np.array([0.0, -Ex / Bz, 0.0])}
    return mesh, species, models, extra


def argon_like_cross_sections(npts=1000):
    """Synthetic electron-argon tables on log-spaced energies 0.01..1000 eV: elastic
    ~1e-19 m^2 with a Ramsauer dip near 0.3 eV, excitation from 11.5 eV peaking at
    2e-21 m^2, ionization from 15.76 eV peaking at 3e-20 m^2 near 100 eV."""
    e = np.logspace(-2, 3, npts)
    el = 1e-19 * (1.0 - 0.9 * np.exp(-(np.log(e / 0.3)) ** 2 / 0.5)) * (1.0 + 0.3 * np.exp(-(np.log(e / 12.0)) ** 2))
    ex = np.where(e > 11.5, 2e-21 * np.clip((e - 11.5) / 8.5, 0, None) * np.exp(1 - np.clip((e - 11.5) / 8.5, 0, None)), 0.0)
    x = np.clip(e / 15.76, 1.0, None)
    iz = np.where(e > 15.76, np.log(x) / x, 0.0)
    iz = 3e-20 * iz / np.max(np.where(e > 15.76, np.log(x) / x, 0.0)) if np.any(iz > 0) else iz
    return e, el, ex, iz


def argon_mcc(ncell=100, ppc=100, T_eV=2.0, Ex=5e4, nN=1e21, TN=300.0, L=None, precision="mixed", dt=None,
              growth=1.5):
    """Electron-impact MCC in an argon-like gas: all-cyclic ncell^3 hex box, ppc electrons
    per cell Maxwellian at T_eV, uniform E (MaxwellSolver none), uniform background.
    ``growth`` sizes the electron capacity (created electrons and ions are appended)."""
    if L is None:
        L = 0.01
    mesh = meshgen.hex_box(ncell, ncell, ncell, L, L, L, cyclic=(True, True, True))
    N = ncell ** 3 * ppc
    mAr = 39.948 * AMU
    e, el, ex, iz = argon_like_cross_sections()
    vmax = np.sqrt(2 * 1000 * QE / ME)
    smax = np.max(np.sqrt(2 * e * QE / ME) * (el + ex + iz))
    if dt is None:
        dt = 0.05 / (nN * smax)
    w = 1.0
    species = [{"name": "e", "mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": int(N * growth)},
               {"name": "Ar+", "mass": mAr, "charge": QE, "nEquivalentParticles": w, "capacity": int(N * (growth - 1)) + 1,
                "solveMovement": False}]
    models = {"ParticlePusher": "Boris", "MaxwellSolver": "none", "ElectronNeutralCollisionModel": "NanbuElectronNeutral",
              "BackgroundGasModel": "uniformBackground", "electronTypeId": 0, "ionTypeId": 1,
              "uniformBackground": {"numberDensity": nN, "temperature": TN, "mass": mAr, "nEquivalentParticles": w},
              "precision": precision}
    extra = {"E": np.tile([Ex, 0.0, 0.0], (ncell ** 3, 1)), "dt": dt, "xs": (e, el, ex, iz), "T_K": T_eV * EV_K,
             "ppc": ppc, "vmax": vmax}
    return mesh, species, models, extra


def poisson_plasma(ncell=100, ppc=8, T_eV=2.0, density=1e16, L=0.01, wp_dt=0.1, guess="previous", tol=1e-8):
    """Field-solve workload on the MCC mesh: ncell^3 hex box of side L, electrons (ppc per
    cell, Maxwellian T_eV) over a neutralising ion background given as spaceCharge, phi = 0
    on the two x walls (electrons leaving there are deleted), cyclic in y and z.
    ElectroStatic PCG+DIC to ||r||/||b|| <= tol every step; dt = wp_dt / omega_p."""
    mesh = meshgen.hex_box(ncell, ncell, ncell, L, L, L, cyclic=(False, True, True))
    ncells = ncell ** 3
    w = density * L ** 3 / (ncells * ppc)
    species = [{"name": "e", "mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": ncells * ppc}]
    models = {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "poissonTolerance": tol,
              "poissonGuess": guess}
    bc = {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 0.0)}
    extra = {"bc": bc, "space_charge": density * QE, "dt": wp_dt / plasma_frequency(density), "T_K": T_eV * EV_K,
             "ppc": ppc, "n_cells": ncells}
    return mesh, species, models, extra


def wall_plasma(ncell=100, ppc=8, T_eV=2.0, density=1e16, L=0.01, wp_dt=0.1, model="MaxwellianThermal",
                wall_T=300.0, tol=1e-8):
    """The field-solve plasma of poisson_plasma with the two x faces as walls at phi = 0:
    electrons reaching them are reflected by the wall model (diffusely at wall_T for
    MaxwellianThermal) and stay in the box."""
    mesh, species, models, extra = poisson_plasma(ncell=ncell, ppc=ppc, T_eV=T_eV, density=density, L=L,
                                                  wp_dt=wp_dt, tol=tol)
    mesh = meshgen.hex_box(ncell, ncell, ncell, L, L, L, cyclic=(False, True, True), boundary_type="wall")
    for p in mesh["patches"]:
        if p["type"] == "wall":
            p["wall_temperature"] = wall_T
    models["WallReflectionModel"] = model
    return mesh, species, models, extra
