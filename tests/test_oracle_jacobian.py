"""Pins of the oracle's analytic Jacobian (oracle/jacobian.c).

The reference builds the Jacobian of finitechem_compute_rhs
(combustion/finitechem.f90:93-175) by CVODE's difference quotients
(FCVDENSE without FCVDENSESETJAC, finitechem.f90:1028); the analytic one
must be the derivative of the same function.  Checked column by column
against Richardson-extrapolated central differences of orc_rhs, which is
itself pinned by the reference mechanisms' golden vectors.  The steps are
scaled by the density (the right-hand side is a polynomial in the
concentrations apart from the temperature and the fall-off blending), so that
trace-species columns are not lost in rounding.  Planted mistakes (the
temperature chain rule dropped, the species rows transposed, one column
scaled by 1 %) must fail the same check.
"""
import numpy as np
import pytest

from tiger_nga_amd import inputs


def central_jacobian(oracle, mech, y, P, frac=1e-4):
    n = len(y)
    R = np.zeros((n, n))
    rho = y[:-1].sum()
    for j in range(n):
        hj = frac * max(abs(y[j]), rho if j < n - 1 else 0.0)

        def D(h):
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            return (oracle.rhs(mech, yp, P) - oracle.rhs(mech, ym, P)) / (2 * h)
        R[:, j] = (4 * D(hj / 2) - D(hj)) / 3
    return R


def column_errors(J, R):
    sp = np.abs(J[:-1] - R[:-1]).max(axis=0) / (np.abs(R[:-1]).max(axis=0) + 1e-300)
    tr = np.abs(J[-1] - R[-1]).max() / np.abs(R[-1]).max()
    return sp.max(), tr


@pytest.mark.parametrize("workload", ["h2", "gri30", "gri30_uneven", "dme39", "ethyred"])
def test_analytic_jacobian_against_central_differences(oracle, workload):
    mech, Y, T, rho, P, dt = inputs.states(workload, n=12, seed=5)
    sol = inputs.packed_state(Y, T, rho)
    checked = 0
    for c in range(len(sol)):
        if abs(T[c] - 1000.0) < 20.0:  # COMPTHERMODATA switches range at 1000 K
            continue
        y = sol[c]
        f, J = oracle.jacobian(mech, y, P)
        f0 = oracle.rhs(mech, y, P)
        assert np.abs(f - f0).max() <= 1e-10 * np.abs(f0).max()
        R = central_jacobian(oracle, mech, y, P)
        e_sp, e_T = column_errors(J, R)
        assert e_sp < 1e-8 and e_T < 1e-8, (c, e_sp, e_T)
        if checked == 0:
            # planted mistakes fail the same check
            rho_c = y[:-1].sum()
            Tc = y[-1] / rho_c
            no_chain = J.copy()
            no_chain[:, :-1] += J[:, -1:] * Tc  # drops df/dT * dT/dy_j = -df/dT T / rho
            transposed = J.copy()
            transposed[:-1, :-1] = J[:-1, :-1].T
            scaled = J.copy()
            scaled[:, 0] *= 1.01
            for bad in (no_chain, transposed, scaled):
                b_sp, b_T = column_errors(bad, R)
                assert max(b_sp, b_T) > 1e-4
        checked += 1
    assert checked >= 8


@pytest.mark.parametrize("mech_wl", ["h2", "gri30_uneven", "dme39"])
def test_dprodrates_low_pressure_branch(oracle, mech_wl):
    """dCDOT/dC and dCDOT/dT against central differences of PRODRATES, on
    generated states and on one with every concentration slightly negative
    (the right-hand side does not clip them, finitechem.f90:121-131), where the
    third-body sums are below zero and the fall-off rates take the M <= 0
    branch, conc = P/RT (GETLINDRATECOEFF, HYDROGENF.f90:295-299)."""
    mech, Y, T, rho, P, dt = inputs.states(mech_wl, n=4, seed=9)
    from tiger_nga_amd.mech import meta
    W = meta.molar_mass(mech)
    for c in range(len(T) + 1):
        if c < len(T):
            C = rho[c] * Y[c] / W
        else:
            c = 0
            C = -1e-7 * (1.0 + np.arange(len(W)) / len(W))
            assert (oracle.prodrates(mech, C, T[c], P)[3][:4] < 0).all()
        cd, dC, dT = oracle.dprodrates(mech, C, T[c], P)
        cd0 = oracle.prodrates(mech, C, T[c], P)[0]
        assert np.allclose(cd, cd0, rtol=1e-12, atol=1e-12 * np.abs(cd0).max())
        n = len(C)
        ctot = C.sum()
        Rc = np.zeros((n, n))
        for j in range(n):
            hj = 1e-4 * max(abs(C[j]), abs(ctot))

            def D(h):
                Cp, Cm = C.copy(), C.copy()
                Cp[j] += h
                Cm[j] -= h
                return (oracle.prodrates(mech, Cp, T[c], P)[0] - oracle.prodrates(mech, Cm, T[c], P)[0]) / (2 * h)
            Rc[:, j] = (4 * D(hj / 2) - D(hj)) / 3
        # column-relative, with a floor relative to each row: an inert
        # species' column (third-body sums only) is many orders below the
        # rows' other entries, where the differences of the full CDOT round
        col = np.abs(Rc).max(axis=0, keepdims=True)
        row = np.abs(Rc).max(axis=1, keepdims=True)
        assert (np.abs(dC - Rc) <= 1e-8 * col + 1e-10 * row).all(), c
        hT = 1e-3 * T[c]
        DT = lambda h: (oracle.prodrates(mech, C, T[c] + h, P)[0] - oracle.prodrates(mech, C, T[c] - h, P)[0]) / (2 * h)
        RT = (4 * DT(hT / 2) - DT(hT)) / 3
        assert np.abs(dT - RT).max() <= 1e-7 * np.abs(RT).max()


def test_analytic_jacobian_integration_matches_dq(oracle):
    """The integrator with the analytic Jacobian converges to the same
    solution as with the difference-quotient one at tight tolerance."""
    for wl in ("h2", "gri30_uneven"):
        mech, Y, T, rho, P, dt = inputs.states(wl, n=16, seed=11)
        sol = inputs.packed_state(Y, T, rho)
        a = oracle.integrate_batch(mech, sol, dt, P, rtol=1e-10, atol=1e-16, nthreads=4, jac=0)
        b, st = oracle.integrate_batch(mech, sol, dt, P, rtol=1e-10, atol=1e-16, nthreads=4, jac=1,
                                       with_stats=True)
        assert (st[:, -1] == 0).all()
        assert (st[:, 2] == 0).all()  # no difference-quotient evaluations
        ya, yb = a[:, :-1] / rho[:, None], b[:, :-1] / rho[:, None]
        assert np.abs(ya - yb).max() / 1.0 < 1e-8
        assert np.abs(a[:, -1] / rho - b[:, -1] / rho).max() < 1e-6
