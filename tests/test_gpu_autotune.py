"""The autotuner measures its candidates on the GPU, records a valid choice
for each knob, and applying the choices gives the step's results bit for bit
as the same Jacobian and step mapping set by hand (the grid and the batched
right-hand-side mapping change scheduling only)."""
import pytest

pytestmark = pytest.mark.gpu


def test_tune_and_apply_h2(tmp_path):
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    from tiger_nga_amd import autotune, surrogate
    from tiger_nga_amd.finitechem import FiniteChem
    rec = autotune.tune("h2", cells=16 ** 3, reps=1, log=print)
    assert rec["rhs_mapping"] in (0, 1) and rec["grid"] >= 0
    assert set(rec["measured"]["rhs_ms"]) == {"lane", "wave"}
    if "surr_nsub" in rec:
        assert 1 <= rec["surr_nsub"] <= 8 and rec["surr_grade"] >= 1.0
        assert rec["surr_adaptive"] in (0, 1) and rec["surr_guess"] in (0, 1) and rec["surr_tol_y"] in autotune.SURR_TOLS
        name = autotune.surr_name(rec["surr_adaptive"], rec["surr_nsub"], rec["surr_grade"], rec["surr_max_att"],
                                  rec["surr_guess"], rec["surr_h0_fac"], rec["surr_max_jac"], rec["surr_keep"],
                                  rec["surr_tol_y"], rec.get("surr_newton_tol"))
        m = rec["measured"]["surrogate"][name]
        assert max(m["max_species_err_generated"], m["max_species_err_relaxed"]) <= autotune.MAX_SPECIES_ERR
    cache = str(tmp_path / "tuning.json")
    autotune.save(rec, 16 ** 3, cache)
    fc = FiniteChem("h2")
    R_d, S0, g, P = surrogate.block_states(fc, "h2", 16 ** 3, 5, 0)
    # the plain step's Jacobian and mapping change the rounding; the grid and
    # the batched right-hand side's mapping do not
    for k in ("jacobian", "step_mapping", "step_waves"):
        if k in rec:
            fc.set_option(k, rec[k])
    ref = S0.clone()
    fc.source_chemistry(R_d, ref, g, 1e-6, P, P)
    done = autotune.apply(fc, 16 ** 3, cache)
    assert done["grid"] == rec["grid"]
    S = S0.clone()
    fc.source_chemistry(R_d, S, g, 1e-6, P, P)
    assert torch.equal(S, ref)
