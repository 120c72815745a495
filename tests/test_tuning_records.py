"""The committed tuning records against the tuner's own selection rules
(host logic only: the records hold the measurements the choices came from)."""
import json

import pytest

from tiger_nga_amd import autotune


def _records():
    with open(autotune.CACHE) as f:
        return json.load(f)


def _screened(m):
    """The candidates within the screening bounds and SURR_MIN_GAIN faster
    than the plain step, fastest first."""
    plain = m["plain_ms"]["generated"] + m["plain_ms"]["relaxed"]
    out = []
    for name, r in m["surrogate"].items():
        if (max(r["max_species_err_generated"], r["max_species_err_relaxed"]) <= autotune.MAX_SPECIES_ERR
                and r["max_T_err_K"] <= autotune.MAX_T_ERR
                and r["ms_generated_plus_relaxed"] * autotune.SURR_MIN_GAIN <= plain):
            out.append((r["ms_generated_plus_relaxed"], name))
    return sorted(out), plain


def test_guided_choice_takes_the_fastest_that_passes_on_the_full_block():
    a = (1.0, 1, 8, 3.0, 64, 0, 16.0, 4, 1.5, 1e-6, 0.03)
    b = (1.2, 1, 8, 3.0, 64, 0, 16.0, 4, 1.5, 5e-7, 0.03)
    name = lambda c: autotune.surr_name(*c[1:])
    assert autotune.guided_choice([a, b], {name(a): {"passed": False}, name(b): {"passed": True}}) == b
    assert autotune.guided_choice([a, b], {name(a): {"passed": True}, name(b): {"passed": True}}) == a
    assert autotune.guided_choice([a, b], {name(a): {"passed": False}, name(b): {"passed": False}}) is None
    assert autotune.guided_choice([], {}) is None


@pytest.mark.parametrize("key", sorted(_records()))
def test_recorded_guided_choice_follows_the_rules(key):
    rec = _records()[key]
    m = rec["measured"]
    if "surrogate" not in m:
        assert "surr_nsub" not in rec  # no network for the mechanism
        return
    screened, plain = _screened(m)
    val = m.get("validation", {}).get("candidates", {})
    # the validated candidates are the fastest screened ones, up to the first
    # that passed
    assert len(val) <= autotune.VALIDATE_K
    assert set(val) == {n for _, n in screened[:len(val)]}
    assert sum(v["passed"] for v in val.values()) <= 1
    for v in val.values():
        assert v["passed"] == (v["max_species_err"] <= autotune.FULL_SPECIES_ERR
                               and v["max_T_err_K"] <= autotune.MAX_T_ERR)
    if "surr_nsub" not in rec:
        # nothing screened, or every validated candidate failed on the full block
        assert all(not v["passed"] for v in val.values())
        assert not screened or len(val) == min(len(screened), autotune.VALIDATE_K)
        return
    name = autotune.surr_name(rec["surr_adaptive"], rec["surr_nsub"], rec["surr_grade"], rec["surr_max_att"],
                              rec["surr_guess"], rec["surr_h0_fac"], rec["surr_max_jac"], rec["surr_keep"],
                              rec["surr_tol_y"], rec.get("surr_newton_tol"))
    assert name in [n for _, n in screened]
    assert val[name]["passed"]
    r = m["surrogate"][name]
    assert m["surrogate_speedup"] == pytest.approx(plain / r["ms_generated_plus_relaxed"])
    # every faster screened candidate was validated and failed
    for t, other in screened:
        if t < r["ms_generated_plus_relaxed"]:
            assert other in val and not val[other]["passed"], other
    assert r["ms_generated_plus_relaxed"] == max(m["surrogate"][n]["ms_generated_plus_relaxed"] for n in val)
