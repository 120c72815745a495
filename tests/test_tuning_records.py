"""The committed tuning records against the tuner's own selection rules
(host logic only: the records hold the measurements the choices came from)."""
import json

import pytest

from tiger_nga_amd import autotune


def _records():
    with open(autotune.CACHE) as f:
        return json.load(f)


def test_guided_choice_needs_the_minimum_gain():
    best = (1.0, 1, 8, 3.0, 64, 0, 16.0, 4, 1.5, 2.5e-7)
    assert autotune.guided_choice(best, 1.0 * autotune.SURR_MIN_GAIN * 1.001) == best
    assert autotune.guided_choice(best, 1.0 * autotune.SURR_MIN_GAIN * 0.999) is None
    assert autotune.guided_choice(None, 10.0) is None


@pytest.mark.parametrize("key", sorted(_records()))
def test_recorded_guided_choice_follows_the_rules(key):
    rec = _records()[key]
    m = rec["measured"]
    if "surr_nsub" not in rec:
        # no guided configuration recorded: either no network for the
        # mechanism, or none within the bounds that pays
        if "surrogate" in m:
            plain = m["plain_ms"]["generated"] + m["plain_ms"]["relaxed"]
            for name, r in m["surrogate"].items():
                ok = (max(r["max_species_err_generated"], r["max_species_err_relaxed"]) <= autotune.MAX_SPECIES_ERR
                      and r["max_T_err_K"] <= autotune.MAX_T_ERR)
                assert not ok or r["ms_generated_plus_relaxed"] * autotune.SURR_MIN_GAIN > plain, name
        return
    name = autotune.surr_name(rec["surr_adaptive"], rec["surr_nsub"], rec["surr_grade"], rec["surr_max_att"],
                              rec["surr_guess"], rec["surr_h0_fac"], rec["surr_max_jac"], rec["surr_keep"],
                              rec["surr_tol_y"], rec.get("surr_newton_tol"))
    r = m["surrogate"][name]
    assert max(r["max_species_err_generated"], r["max_species_err_relaxed"]) <= autotune.MAX_SPECIES_ERR
    assert r["max_T_err_K"] <= autotune.MAX_T_ERR
    plain = m["plain_ms"]["generated"] + m["plain_ms"]["relaxed"]
    assert r["ms_generated_plus_relaxed"] * autotune.SURR_MIN_GAIN <= plain
    assert m["surrogate_speedup"] == pytest.approx(plain / r["ms_generated_plus_relaxed"])
    # the fastest of the configurations within the bounds
    for other, o in m["surrogate"].items():
        if (max(o["max_species_err_generated"], o["max_species_err_relaxed"]) <= autotune.MAX_SPECIES_ERR
                and o["max_T_err_K"] <= autotune.MAX_T_ERR):
            assert o["ms_generated_plus_relaxed"] >= r["ms_generated_plus_relaxed"], other
