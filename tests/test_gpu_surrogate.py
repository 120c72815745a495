"""The surrogate-guided step through the C ABI, with the committed trained
networks (tiger_nga_amd/models/): every cell ends either accepted by the
guided solve (then within the acceptance bound of the plain FP64 result and
element-conserving) or bit-identical to the plain step, which integrated it;
and what the trained network gains over the same solve without it."""
import os

import numpy as np
import pytest

from tiger_nga_amd.mech import meta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    return torch


def _model(mech):
    from tiger_nga_amd import surrogate
    path = surrogate.model_path(mech)
    assert os.path.exists(path), "trained network missing: %s" % path
    return surrogate.Surrogate.load(path)


@pytest.mark.parametrize("mech,workload,n,relax,adaptive,guess,split,h0f", [
    ("h2", "h2", 16 ** 3, 0, 1, 0, 1, 0.0), ("h2", "h2", 16 ** 3, 1, 1, 0, 1, 0.0),
    ("gri30", "gri30", 16 ** 3, 0, 1, 0, 1, 0.0), ("gri30", "gri30", 16 ** 3, 1, 1, 0, 1, 0.0),
    ("gri30", "gri30", 16 ** 3, 0, 1, 1, 1, 0.0), ("gri30", "gri30", 16 ** 3, 1, 0, 1, 1, 0.0),
    ("gri30", "gri30", 16 ** 3, 0, 1, 0, 0, 0.0), ("gri30", "gri30", 16 ** 3, 0, 1, 0, 1, 16.0),
    ("gri30", "gri30", 16 ** 3, 1, 1, 0, 1, 16.0), ("ethyred", "ethyred", 16 ** 3, 0, 1, 0, 1, 0.0),
    ("dme39", "dme39", 16 ** 3, 0, 1, 0, 1, 0.0)])
def test_cells_accepted_within_bound_or_plain(torch_dev, mech, workload, n, relax, adaptive, guess, split, h0f):
    """adaptive: error-controlled (1) or fixed (0) schedule; guess: Newton
    from the network's prediction (1) or without the network (0); split:
    the Jacobian with the rate constants evaluated once (1) or every column a
    full difference quotient (0); h0f: first step from the initial rate."""
    torch = torch_dev
    from tiger_nga_amd import surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    fc = FiniteChem(mech, 1e-6, 1e-12)
    R_d, S0, g, P = surrogate.block_states(fc, workload, n, 3, relax)
    dt = {"h2": 1e-6, "gri30": 5e-7, "ethyred": 5e-7, "dme39": 2e-7}[mech]
    nsp = fc.nsp
    Sp = S0.clone()
    HRp = torch.zeros_like(R_d)
    stp = fc.source_chemistry(R_d, Sp, g, dt, P, P, HR_gp=HRp)
    # the exact solution to within 1e-9: the plain step at rtol 1e-10
    ft = FiniteChem(mech, 1e-10, 1e-16)
    ft.set_option("max_steps", 200000)
    St = S0.clone()
    ft.source_chemistry(R_d, St, g, dt, P, P)
    ft.close()
    sur = _model(mech)
    sur.attach(fc)
    fc.set_option("surr_adaptive", adaptive)
    fc.set_option("surr_guess", guess)
    fc.set_option("surr_split_jac", split)
    fc.set_option("surr_h0_fac", h0f)
    if not adaptive:
        fc.set_option("surr_nsub", 4)
        fc.set_option("surr_grade", 1.0)
        fc.set_option("surr_tol_y", 1e-6)
    Ss = S0.clone()
    HRs = torch.zeros_like(R_d)
    st, ss = fc.source_chemistry_surrogate(R_d, Ss, g, dt, P, P, HR_gp=HRs)
    assert ss["accepted"] + ss["rejected"] == n == st["cells"]
    assert st["failed"] == 0 and stp["failed"] == 0
    same = (Ss == Sp).all(dim=3)
    assert int((~same).sum()) <= ss["accepted"]
    # accepted cells: within 1e-6 of the exact solution in every species (or
    # no further from it than the plain step is), within 2e-2 K of the plain T
    dY = (Ss[..., :nsp] - Sp[..., :nsp]).abs().amax(dim=3)
    dT = (Ss[..., nsp] - Sp[..., nsp]).abs()
    ex = float((Ss[..., :nsp] - St[..., :nsp]).abs().max())
    ex_plain = float((Sp[..., :nsp] - St[..., :nsp]).abs().max())
    print("\nmax |dY| against the exact solution: guided %.2e, plain %.2e" % (ex, ex_plain))
    assert ex <= max(1e-6, ex_plain) and float(dT.max()) <= 2e-2
    # element conservation of every cell (Y in mass fractions)
    Et = torch.from_numpy(np.ascontiguousarray((meta.element_matrix(mech) / meta.molar_mass(mech)[None, :]).T)).cuda()
    el0 = S0[..., :nsp].contiguous() @ Et
    el1 = Ss[..., :nsp].contiguous() @ Et
    elp = Sp[..., :nsp].contiguous() @ Et
    cons = float(((el1 - el0).abs() / el0.abs().clamp_min(1e-30)).max())
    cons_plain = float(((elp - el0).abs() / el0.abs().clamp_min(1e-30)).max())
    print("\nelement conservation: guided %.2e, plain %.2e (max relative change)" % (cons, cons_plain))
    assert cons <= max(1e-12, 2.0 * cons_plain)
    print("\n%s relax %d adaptive %d guess %d split %d h0f %g: accepted %d / %d, rejected (Newton %d, error %d, "
          "enthalpy %d); max|dY| %.2e max|dT| %.2e K" % (mech, relax, adaptive, guess, split, h0f, ss["accepted"], n,
                                                         ss["rej_newton"],
                                           ss["rej_error"], ss["rej_enthalpy"], float(dY.max()), float(dT.max())))
    assert ss["substeps"] >= ss["accepted"]
    if adaptive and relax:
        assert ss["accepted"] >= 0.9 * n


def test_network_guess_gain(torch_dev):
    """Newton iterations of the fixed-schedule guided solve (4 equal
    sub-steps, stages converged to 1e-8 of the acceptance tolerance,
    acceptance off) with the network's guesses against the same solve started
    from the previous stage value, on relaxed GRI-3.0 states: the trained
    network saves iterations there.  (In the adaptive schedule the derivative
    start is the better one; test_network_guess_adaptive records both.)"""
    from tiger_nga_amd import surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    fc = FiniteChem("gri30", 1e-6, 1e-12)
    R_d, S0, g, P = surrogate.block_states(fc, "gri30", 32 ** 3, 4, 1)
    _model("gri30").attach(fc)
    fc.set_option("surr_adaptive", 0)
    fc.set_option("surr_nsub", 4)
    fc.set_option("surr_grade", 1.0)
    fc.set_option("surr_newton_max", 20)
    fc.set_option("surr_tol_y", 1.0)
    fc.set_option("surr_tol_t", 1e4)
    fc.set_option("surr_newton_tol", 1e-8)
    it = {}
    for guess in (1, 0):
        fc.set_option("surr_guess", guess)
        S = S0.clone()
        _, ss = fc.source_chemistry_surrogate(R_d, S, g, 5e-7, P, P)
        it[guess] = ss["newton_iters"] / (32 ** 3)
    print("\nNewton iterations per cell: with the network %.2f, without %.2f (%.1f %% fewer)"
          % (it[1], it[0], 100.0 * (1.0 - it[1] / it[0])))
    assert it[1] < it[0]


def test_network_guess_adaptive(torch_dev):
    """The adaptive schedule on raw GRI-3.0 states with and without the
    network's guesses: both end within 1e-6 of the plain step on every cell;
    the iteration counts of both are recorded (the derivative start needs
    fewer on the benchmark states, which is why it is the default)."""
    from tiger_nga_amd import surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    fc = FiniteChem("gri30", 1e-6, 1e-12)
    R_d, S0, g, P = surrogate.block_states(fc, "gri30", 16 ** 3, 6, 0)
    nsp = fc.nsp
    Sp = S0.clone()
    fc.source_chemistry(R_d, Sp, g, 5e-7, P, P)
    _model("gri30").attach(fc)
    fc.set_option("surr_probe", 0)
    it = {}
    for guess in (1, 0):
        fc.set_option("surr_guess", guess)
        S = S0.clone()
        _, ss = fc.source_chemistry_surrogate(R_d, S, g, 5e-7, P, P)
        assert float((S[..., :nsp] - Sp[..., :nsp]).abs().max()) <= 1e-6
        it[guess] = (ss["newton_iters"] / (16 ** 3), ss["accepted"] / (16 ** 3))
    print("\nadaptive schedule, Newton iterations per cell (accepted fraction): with the network %.1f (%.3f), "
          "derivative start %.1f (%.3f)" % (it[1][0], it[1][1], it[0][0], it[0][1]))


@pytest.mark.parametrize("mech,workload,relax", [("gri30", "gri30_uneven", 0), ("gri30", "gri30_uneven", 1),
                                                 ("h2", "h2", 1), ("dme39", "dme39", 1)])
def test_standalone_surrogate(torch_dev, mech, workload, relax):
    """surr_mode 1, the standalone surrogate: the network's predictions at the
    stage times of 4 equal sub-steps are the stage values, each checked by one
    FP64 Newton correction of its stage equation (accepted within surr_sa_tol
    = 1 acceptance tolerance), then the embedded error estimate, the element
    projection and the enthalpy check; the plain FP64 step integrates every
    other cell.  Every cell ends accepted (within 4e-6 of the tight solution --
    four sub-steps, each within 1e-6 -- and element-conserving) or bit-identical
    to the plain step.  The accepted fraction with the network is printed next
    to the same check on the zero-change prediction (the method without the
    network)."""
    torch = torch_dev
    from tiger_nga_amd import inputs, surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    n = 16 ** 3
    fc = FiniteChem(mech, 1e-6, 1e-12)
    R_d, S0, g, P = surrogate.block_states(fc, workload, n, 3, relax)
    dt = inputs.WORKLOADS[workload][3]
    nsp = fc.nsp
    Sp = S0.clone()
    fc.source_chemistry(R_d, Sp, g, dt, P, P)
    ft = FiniteChem(mech, 1e-10, 1e-16, 200000)
    St = S0.clone()
    ft.source_chemistry(R_d, St, g, dt, P, P)
    ft.close()
    _model(mech).attach(fc)
    fc.set_option("surr_probe", 0)
    fc.set_option("surr_mode", 1)
    fc.set_option("surr_nsub", 4)
    fc.set_option("surr_grade", 1.0)
    fc.set_option("surr_tol_y", 1e-6)
    acc = {}
    Et = torch.from_numpy(np.ascontiguousarray((meta.element_matrix(mech) / meta.molar_mass(mech)[None, :]).T)).cuda()
    el0 = S0[..., :nsp].contiguous() @ Et
    for guess in (1, 0):
        fc.set_option("surr_guess", guess)
        Ss = S0.clone()
        st, ss = fc.source_chemistry_surrogate(R_d, Ss, g, dt, P, P)
        assert ss["accepted"] + ss["rejected"] == n == st["cells"] and st["failed"] == 0
        same = (Ss == Sp).all(dim=3)
        assert int((~same).sum()) <= ss["accepted"]
        sel = ~same
        err = float((Ss[..., :nsp] - St[..., :nsp]).abs().amax(dim=3)[sel].max()) if bool(sel.any()) else 0.0
        el1 = Ss[..., :nsp].contiguous() @ Et
        cons = float(((el1 - el0).abs() / el0.abs().clamp_min(1e-30))[sel].max()) if bool(sel.any()) else 0.0
        acc[guess] = (ss["accepted"] / n, err, ss["rej_newton"], ss["rej_error"], ss["rej_enthalpy"])
        print("\n%s relax %d standalone, %s: accepted %.4f, max |dY| of accepted cells vs the tight solution %.2e, "
              "element drift %.1e; rejected: residual %d, error estimate %d, enthalpy %d"
              % (workload, relax, "network" if guess else "zero-change prediction", acc[guess][0], err, cons,
                 ss["rej_newton"], ss["rej_error"], ss["rej_enthalpy"]))
        assert err <= 4e-6 and cons <= 1e-12
    fc.close()


def test_fallback_through_thread_per_cell_step(torch_dev):
    """The guided call with the plain step in its thread-per-cell mapping
    (step_mapping 1, H2): the rejected cells go through that kernel, whose
    workspace the call sizes before its first launch; every cell ends either
    accepted or bit-identical to the thread-per-cell plain step."""
    torch = torch_dev
    from tiger_nga_amd import surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    fc = FiniteChem("h2", 1e-6, 1e-12)
    fc.set_option("step_mapping", 1)
    R_d, S0, g, P = surrogate.block_states(fc, "h2", 16 ** 3, 3, 0)
    Sp = S0.clone()
    stp = fc.source_chemistry(R_d, Sp, g, 1e-6, P, P)
    _model("h2").attach(fc)
    fc.set_option("surr_probe", 0)
    Ss = S0.clone()
    st, ss = fc.source_chemistry_surrogate(R_d, Ss, g, 1e-6, P, P)
    fc.close()
    print("\nh2 guided with the thread-per-cell fallback: accepted %d, rejected %d" % (ss["accepted"], ss["rejected"]))
    assert stp["failed"] == 0 and st["failed"] == 0 and ss["accepted"] + ss["rejected"] == 16 ** 3
    same = (Ss == Sp).all(dim=3)
    assert ss["rejected"] > 0 and int((~same).sum()) <= ss["accepted"]
