"""Seeded synthetic chemistry states for the benchmark workloads.

The reference seeds its RNG from the clock (library/random.f90:100-116) and
ships no case data, so the states are synthetic, generated per cell from a
seeded numpy PCG64 stream:

* ``h2``      9 species, Z~U(0,1) between H2 fuel and air (Y_O2 0.233,
              Y_N2 0.767); progress c~U(0,1) converts c of the limiting
              reactant to H2O (2 H2 + O2 -> 2 H2O); H, O, OH, HO2, H2O2
              log-uniform in [1e-6, 1e-3]; renormalised; T~U(900, 2200) K.
* ``gri30``   36 species, Z~U(0, 0.15) between CH4 and air; c~U(0,1) of
              CH4 + 2 O2 burnt, carbon to 90% CO2 / 10% CO, hydrogen to 95%
              H2O / 5% H2; all other species log-uniform in [1e-8, 1e-3];
              T~U(800, 2400) K.
* ``gri30_uneven``  as gri30 with 70% near-equilibrium cells (T 1800-2300 K,
              c 0.9-1), 20% ignition-front cells (T 1000-1300 K, c 0.05-0.2)
              and 10% cold mixing cells (T 300-600 K, c 0-0.05), randomly
              interleaved.
* ``dme39``   39 species, as gri30 with CH3OCH3 fuel (CH3OCH3 + 3 O2 ->
              2 CO2 + 3 H2O at full conversion), T~U(600, 2200) K.
* ``ethyred`` 45 species (EthyRedF, the mechanism the reference links by
              default), as gri30 with C2H4 fuel (C2H4 + 3 O2 -> 2 CO2 +
              2 H2O), 64^3 cells, 1 atm, dt 5e-7 s, T~U(800, 2400) K.

Density follows the EOS of finitechem_density (combustion/finitechem.f90:
324-345): rho = P W_mix / (R T), R = 8314.34 J/kmol/K.
"""
from __future__ import annotations

import numpy as np

from .mech import meta

R_CST = 8314.34
P_ATM = 101325.0

WORKLOADS = {
    # name: (mechanism, cells, pressure, dt)
    "h2": ("h2", 32 ** 3, P_ATM, 1e-6),
    "gri30": ("gri30", 64 ** 3, P_ATM, 5e-7),
    "gri30_uneven": ("gri30", 128 ** 3, P_ATM, 5e-7),
    "dme39": ("dme39", 128 ** 3, 10 * P_ATM, 2e-7),
    "ethyred": ("ethyred", 64 ** 3, P_ATM, 5e-7),
}


def _mix(mech, n, Zmax, fuel, rng):
    sp = meta.species(mech)
    Y = np.zeros((n, len(sp)))
    Z = rng.uniform(0.0, Zmax, n)
    Y[:, sp.index(fuel)] = Z
    Y[:, sp.index("O2")] = (1.0 - Z) * 0.233
    Y[:, sp.index("N2")] = (1.0 - Z) * 0.767
    return Y


def _burn(mech, Y, c, fuel, nC, nH, o2_per_fuel, co_frac, h2_frac):
    """Convert c of the limiting reactant of fuel + o2_per_fuel O2."""
    sp, W = meta.species(mech), meta.molar_mass(mech)
    i_f, i_o2 = sp.index(fuel), sp.index("O2")
    nf = Y[:, i_f] / W[i_f]
    no2 = Y[:, i_o2] / W[i_o2]
    xi = c * np.minimum(nf, no2 / o2_per_fuel)
    carbon, hyd = nC * xi, nH * xi
    co2 = carbon * (1.0 - co_frac)
    co = carbon * co_frac
    h2o = 0.5 * hyd * (1.0 - h2_frac)
    h2 = 0.5 * hyd * h2_frac
    o_fuel = {"CH4": 0, "CH3OCH3": 1}.get(fuel, 0) * xi
    o2_used = 0.5 * (2 * co2 + co + h2o - o_fuel)
    Y[:, i_f] -= xi * W[i_f]
    Y[:, i_o2] -= o2_used * W[i_o2]
    for name, mol in (("CO2", co2), ("CO", co), ("H2O", h2o), ("H2", h2)):
        if name in sp and np.any(mol):
            Y[:, sp.index(name)] += mol * W[sp.index(name)]
    return Y


def _minors(mech, Y, lo, hi, majors, rng):
    sp = meta.species(mech)
    for k, s in enumerate(sp):
        if s not in majors:
            Y[:, k] = 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), Y.shape[0])
    Y = np.maximum(Y, 0.0)
    return Y / Y.sum(axis=1, keepdims=True)


def states(workload: str, n: int | None = None, seed: int = 0):
    """Return (mech, Y[n, nsp], T[n], rho[n], P, dt) for a workload."""
    mech, ncell, P, dt = WORKLOADS[workload]
    n = ncell if n is None else n
    rng = np.random.Generator(np.random.PCG64(seed))
    if mech == "h2":
        sp, W = meta.species(mech), meta.molar_mass(mech)
        Z = rng.uniform(0.0, 1.0, n)
        Y = np.zeros((n, len(sp)))
        Y[:, sp.index("H2")] = Z
        Y[:, sp.index("O2")] = (1 - Z) * 0.233
        Y[:, sp.index("N2")] = (1 - Z) * 0.767
        c = rng.uniform(0.0, 1.0, n)
        iH2, iO2, iH2O = sp.index("H2"), sp.index("O2"), sp.index("H2O")
        xi = c * np.minimum(Y[:, iH2] / W[iH2] / 2.0, Y[:, iO2] / W[iO2])
        Y[:, iH2] -= 2 * xi * W[iH2]
        Y[:, iO2] -= xi * W[iO2]
        Y[:, iH2O] += 2 * xi * W[iH2O]
        Y = _minors(mech, Y, 1e-6, 1e-3, ("N2", "O2", "H2", "H2O"), rng)
        T = rng.uniform(900.0, 2200.0, n)
    else:
        fuel = {"gri30": "CH4", "dme39": "CH3OCH3", "ethyred": "C2H4"}[mech]
        nC, nH, o2 = {"CH4": (1, 4, 2.0), "CH3OCH3": (2, 6, 3.0), "C2H4": (2, 4, 3.0)}[fuel]
        Y = _mix(mech, n, 0.15, fuel, rng)
        if workload == "gri30_uneven":
            u = rng.random(n)
            kind = np.where(u < 0.7, 0, np.where(u < 0.9, 1, 2))
            c = np.choose(kind, [rng.uniform(0.9, 1.0, n), rng.uniform(0.05, 0.2, n), rng.uniform(0.0, 0.05, n)])
            T = np.choose(kind, [rng.uniform(1800, 2300, n), rng.uniform(1000, 1300, n), rng.uniform(300, 600, n)])
        else:
            c = rng.uniform(0.0, 1.0, n)
            T = rng.uniform(600.0, 2200.0, n) if mech == "dme39" else rng.uniform(800.0, 2400.0, n)
        Y = _burn(mech, Y, c, fuel, nC, nH, o2, 0.1, 0.05)
        Y = _minors(mech, Y, 1e-8, 1e-3, ("N2", "O2", fuel, "CO2", "H2O", "CO", "H2"), rng)
    W = meta.molar_mass(mech)
    Wmix = 1.0 / (Y / W).sum(axis=1)
    rho = P * Wmix / (R_CST * T)
    return mech, Y, T, rho, P, dt


def packed_state(Y, T, rho):
    """sol[n, neq] = [rho*Y, rho*T] as finitechem.f90:1002-1003 packs it."""
    return np.concatenate([np.maximum(rho[:, None] * Y, 0.0), (rho * T)[:, None]], axis=1)


def fortran_fields(Y, T, rho, nx, ny, nz, nover=0, extra_scalars=0):
    """RHO_(nxo,nyo,nzo), SC_(..., nscalar) Fortran-ordered, species first
    then T (isc_sc = 1, isc_T = nsp + 1), with ghost layers filled with the
    nearest interior value; returns (RHO_, SC_, isc_sc, isc_T)."""
    n, nsp = Y.shape
    assert n == nx * ny * nz
    nscalar = nsp + 1 + extra_scalars
    sh = (nx + 2 * nover, ny + 2 * nover, nz + 2 * nover)
    RHO_ = np.zeros(sh, order="F")
    SC_ = np.zeros(sh + (nscalar,), order="F")
    core = (slice(nover, nover + nx), slice(nover, nover + ny), slice(nover, nover + nz))
    RHO_[core] = rho.reshape((nx, ny, nz), order="F")
    for k in range(nsp):
        SC_[core + (k,)] = Y[:, k].reshape((nx, ny, nz), order="F")
    SC_[core + (nsp,)] = T.reshape((nx, ny, nz), order="F")
    if nover:
        RHO_[:] = np.pad(RHO_[core], nover, mode="edge")
        for s in range(nsp + 1):
            SC_[..., s] = np.pad(SC_[core + (s,)], nover, mode="edge")
    return RHO_, SC_, 1, nsp + 1
