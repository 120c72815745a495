"""The analytic Jacobian (option "jacobian" 1) on the GPU against the CPU
oracle.

The reference builds the Jacobian by CVODE's difference quotients (FCVDENSE,
combustion/finitechem.f90:1028); the analytic one is the exact derivative of
the same right-hand side (pinned by central differences in
tests/test_oracle_jacobian.py, the device code against the oracle's on the
CPU wave emulator in tests/test_wave_emu.py).  Here the step kernel's
analytic instance runs finitechem_source_chemistry on the benchmark
workloads' states and must match (a) the oracle integrating with the same
analytic Jacobian (same CVODE-semantics decisions) and (b) the oracle with
the reference's difference-quotient Jacobian, both at tight tolerance, to
rel 1e-8 in Y where rtol governs and 1e-6 K in T; with no difference-
quotient evaluation counted.  The surrogate-guided solve with the analytic
Jacobian is held to the same bound against the tight oracle as with the
difference-quotient one.
"""
import os

import numpy as np
import pytest

from tiger_nga_amd import inputs
from tiger_nga_amd.finitechem import Grid

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    return torch


def _block(workload, nx, seed):
    mech, Y, T, rho, P, dt = inputs.states(workload, n=nx ** 3, seed=seed)
    RHO_, SC_, isc_sc, isc_T = inputs.fortran_fields(Y, T, rho, nx, nx, nx, nover=1)
    g = Grid.block(nx, nx, nx, SC_.shape[3], isc_sc, isc_T, nover=1)
    return mech, Y, T, rho, P, dt, RHO_, SC_, g


def _interior(SC_, nsp):
    core = SC_[1:-1, 1:-1, 1:-1, :nsp + 1]
    return np.stack([core[..., k].ravel(order="F") for k in range(nsp + 1)], axis=1)


@pytest.mark.parametrize("workload,rtol,atol", [("h2", 1e-10, 1e-16), ("gri30_uneven", 1e-10, 1e-16),
                                                ("dme39", 1e-12, 1e-20), ("ethyred", 1e-10, 1e-16)])
def test_step_analytic_jacobian_vs_oracle(workload, rtol, atol, oracle, torch_dev):
    from tiger_nga_amd.finitechem import FiniteChem
    nx = 5
    mech, Y, T, rho, P, dt, RHO_, SC_, g = _block(workload, nx, seed=21)
    nsp = Y.shape[1]
    fc = FiniteChem(mech, rtol, atol, 200000)
    fc.set_option("jacobian", 1)
    S = SC_.copy(order="F")
    st = fc.source_chemistry(RHO_, S, g, dt, P, P)
    fc.close()
    got = _interior(S, nsp)
    sol = inputs.packed_state(Y, T, rho)
    res = {}
    for jac in (1, 0):
        out, ost = oracle.integrate_batch(mech, sol, dt, P, rtol=rtol, atol=atol, max_steps=200000,
                                          nthreads=THREADS, with_stats=True, jac=jac)
        assert (ost[:, -1] == 0).all()
        ref = np.concatenate([out[:, :-1] / rho[:, None], (out[:, -1] / rho)[:, None]], axis=1)
        rel = np.abs(got[:, :nsp] - ref[:, :nsp]) / np.maximum(np.abs(ref[:, :nsp]), 1e-12)
        dom = rtol * rho[:, None] * np.abs(ref[:, :nsp]) >= 10.0 * atol
        res[jac] = (rel[dom].max(), np.abs(got[:, nsp] - ref[:, nsp]).max(), int(ost[:, 0].sum()),
                    int(ost[:, 3].sum()))
    print("\n%s analytic J: gpu steps %d, nje %d, nfe %d; vs oracle analytic: rel Y %.2e, |dT| %.2e K, steps %d; "
          "vs oracle difference quotients: rel Y %.2e, |dT| %.2e K, steps %d"
          % (workload, st["nst"], st["nje"], st["nfe"], res[1][0], res[1][1], res[1][2], res[0][0], res[0][1],
             res[0][2]))
    assert st["failed"] == 0 and st["nfe_dq"] == 0 and st["nje"] > 0
    assert abs(st["nst"] - res[1][2]) <= 0.01 * res[1][2]
    for jac in (1, 0):
        assert res[jac][0] <= 1e-8 and res[jac][1] <= 1e-6, (jac, res[jac])


def test_step_analytic_jacobian_reference_tolerance(oracle, torch_dev):
    """At the reference tolerances (rtol 1e-6, atol 1e-12) the analytic and
    difference-quotient steps are both within the tolerance band of the tight
    oracle; the analytic one evaluates no difference quotients."""
    from tiger_nga_amd.finitechem import FiniteChem
    nx = 6
    mech, Y, T, rho, P, dt, RHO_, SC_, g = _block("gri30_uneven", nx, seed=22)
    nsp = Y.shape[1]
    out = oracle.integrate_batch(mech, inputs.packed_state(Y, T, rho), dt, P, rtol=1e-10, atol=1e-16,
                                 max_steps=200000, nthreads=THREADS)
    ref = out[:, :-1] / rho[:, None]
    errs = {}
    for jac in (0, 1):
        fc = FiniteChem(mech, 1e-6, 1e-12)
        fc.set_option("jacobian", jac)
        S = SC_.copy(order="F")
        st = fc.source_chemistry(RHO_, S, g, dt, P, P)
        fc.close()
        assert st["failed"] == 0
        if jac:
            assert st["nfe_dq"] == 0
        errs[jac] = (np.abs(_interior(S, nsp)[:, :nsp] - ref).max(), st["nst"], st["nfe"] + st["nfe_dq"])
    print("\nrtol 1e-6: difference quotients max |dY| %.2e (steps %d, RHS %d); analytic %.2e (steps %d, RHS %d)"
          % (errs[0] + errs[1]))
    assert errs[0][0] <= 1e-6 and errs[1][0] <= 1e-6


def test_guided_analytic_jacobian_vs_tight_oracle(oracle, torch_dev):
    """The surrogate-guided solve with the analytic Jacobian on the uneven
    GRI-3.0 states: accepted cells and results held to the same bound against
    the tight oracle as the difference-quotient guided solve."""
    from tiger_nga_amd import autotune, surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    nx = 8
    mech, Y, T, rho, P, dt, RHO_, SC_, g = _block("gri30_uneven", nx, seed=5)
    nsp = Y.shape[1]
    out = oracle.integrate_batch(mech, inputs.packed_state(Y, T, rho), dt, P, rtol=1e-10, atol=1e-16,
                                 max_steps=200000, nthreads=THREADS)
    ref = out[:, :-1] / rho[:, None]
    errs = {}
    for jac in (0, 1):
        fc = FiniteChem(mech, 1e-6, 1e-12)
        autotune.apply(fc, inputs.WORKLOADS["gri30_uneven"][1])
        fc.set_option("jacobian", jac)
        fc.load_surrogate(surrogate.model_path(mech))
        fc.set_option("surr_probe", 0)
        S = SC_.copy(order="F")
        st, ss = fc.source_chemistry_surrogate(RHO_, S, g, dt, P, P)
        fc.close()
        assert st["failed"] == 0 and ss["accepted"] + ss["rejected"] == nx ** 3
        errs[jac] = (np.abs(_interior(S, nsp)[:, :nsp] - ref).max(), ss["accepted"], ss["newton_iters"])
    print("\nguided: difference-quotient J max |dY| %.2e (accepted %d, Newton %d); analytic J %.2e (accepted %d, "
          "Newton %d)" % (errs[0] + errs[1]))
    assert errs[1][1] >= 0.9 * errs[0][1]
    assert errs[1][0] <= max(1e-6, 1.5 * errs[0][0])
