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
MECHS = ("h2", "gri30", "dme39", "ethyred")


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



def _ref_lib(mech):
    path = os.path.join(os.path.dirname(GOLD), "..", "oracle", "_ref", "libref_%s.so" % mech)
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference mechanism sources absent)")
    import ctypes
    return ctypes.CDLL(os.path.abspath(path), mode=os.RTLD_LOCAL)


@pytest.mark.parametrize("mech", MECHS)
def test_oracle_vs_reference_mechanism_random_states(mech, oracle):
    """Beyond the committed vectors: 500 fresh states per mechanism through the
    reference mechanism file itself (oracle/_ref, compiled from the reference
    sources by oracle/Makefile) and the oracle, K and W bit for bit."""
    import ctypes
    lib = _ref_lib(mech)
    dp = ctypes.POINTER(ctypes.c_double)
    d = _gold(mech)
    nsp, nslot = len(d["names"]), int(d["nslot"])
    rng = np.random.default_rng(99)
    for s in range(500):
        T = rng.uniform(280.0, 3200.0)
        P = rng.choice([101325.0, 1013250.0, 5e4])
        C = 10.0 ** rng.uniform(-14.0, -1.0, nsp)
        C[rng.random(nsp) < 0.1] = 0.0
        cd, w, k, m = np.zeros(nsp), np.zeros(nslot), np.zeros(nslot), np.zeros(64)
        c = C.copy()
        lib.prodrates_(cd.ctypes.data_as(dp), w.ctypes.data_as(dp), k.ctypes.data_as(dp), c.ctypes.data_as(dp),
                       m.ctypes.data_as(dp), ctypes.byref(ctypes.c_double(T)), ctypes.byref(ctypes.c_double(P)))
        cd_o, K_o, W_o, _ = oracle.prodrates(mech, C, T, P, nslot)
        np.testing.assert_array_equal(K_o[:nslot], k)
        np.testing.assert_array_equal(W_o[:nslot], w)
        assert np.abs(cd_o - cd).max() <= 1e-14 * max(np.abs(w).sum(), 1e-300)
