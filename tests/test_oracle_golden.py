"""The CPU oracle pinned against the reference's own mechanism routines.

Golden vectors (tests/golden/*_mech.npz) are outputs of PRODRATES and
COMPTHERMODATA from combustion/HYDROGENF.f90, grimech30F.f90 and DME39F.f90
compiled unchanged with flang (tests/golden/make_golden.py).  The oracle's
table-driven restatement must reproduce every rate coefficient K and rate of
progress W bit for bit (same float32-widened literals, same expression
order) and CDOT to rounding of the regrouped sums."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MECHS = ("h2", "gri30", "dme39")


def _gold(mech):
    return np.load(os.path.join(GOLD, "%s_mech.npz" % mech))


@pytest.mark.parametrize("mech", MECHS)
def test_rate_coefficients_bit_exact(mech, oracle):
    d = _gold(mech)
    nslot = int(d["nslot"])
    for s in range(len(d["T"])):
        cd, K, W, M = oracle.prodrates(mech, d["C"][s], d["T"][s], d["P"][s], nslot)
        np.testing.assert_array_equal(K[:nslot], d["K"][s])
        np.testing.assert_array_equal(W[:nslot], d["W"][s])
        scale = np.abs(d["W"][s]).sum()
        assert np.abs(cd - d["CDOT"][s]).max() <= 1e-14 * max(scale, 1e-300)


@pytest.mark.parametrize("mech", MECHS)
def test_thermo_bit_exact(mech, oracle):
    d = _gold(mech)
    nsp = len(d["names"])
    for s, T in enumerate(d["T_thermo"]):
        h, cp = oracle.thermo(mech, T, nsp)
        np.testing.assert_array_equal(h, d["H"][s])
        np.testing.assert_array_equal(cp, d["CP"][s])


@pytest.mark.parametrize("mech", MECHS)
def test_species_and_molar_mass(mech):
    from tiger_nga_amd.mech import meta
    d = _gold(mech)
    assert tuple(d["names"]) == meta.species(mech)
    np.testing.assert_array_equal(meta.molar_mass(mech), d["W_mol"])


@pytest.mark.parametrize("mech", MECHS)
def test_enthalpy_derivative_is_cp(mech, oracle):
    """dh/dT = cp of the NASA-7 pairs (a closed form of the table, not a
    restatement): central difference on both ranges."""
    nsp = len(_gold(mech)["names"])
    for T in (500.0, 900.0, 1500.0, 2500.0):
        hp, _ = oracle.thermo(mech, T + 1e-3, nsp)
        hm, _ = oracle.thermo(mech, T - 1e-3, nsp)
        _, cp = oracle.thermo(mech, T, nsp)
        np.testing.assert_allclose((hp - hm) / 2e-3, cp, rtol=1e-6)


@pytest.mark.parametrize("mech", MECHS)
def test_mass_and_element_conservation_of_rates(mech, oracle):
    """sum_k W_k CDOT_k = 0 and sum_k E_ek CDOT_k = 0 for every state: a
    dropped or mis-signed stoichiometric term breaks one of them."""
    from tiger_nga_amd.mech import meta
    d = _gold(mech)
    W = meta.molar_mass(mech)
    E = meta.element_matrix(mech)
    for s in range(len(d["T"])):
        scale = np.abs(d["W"][s]).sum() * W.max()
        cd_o, _, Wsl, _ = oracle.prodrates(mech, d["C"][s], d["T"][s], d["P"][s])
        assert abs((W * cd_o).sum()) <= 1e-6 * max(scale, 1e-300)   # binary32 molar masses
        assert np.abs(E @ cd_o).max() <= 1e-12 * max(np.abs(Wsl).sum() * 6, 1e-300)

