"""Autotuner: per mechanism and device, time the mapping and launch choices on
a sample of the actual states and keep the fastest.

Choices (all exposed as context options of include/tiger_chem.h):

* ``rhs_mapping``  batched right-hand side (tgr_chem_rhs_device): one state
  per lane (generated straight-line code) or one state per wavefront;
* ``step_mapping`` the plain step: a wavefront per cell or, for the
  mechanisms small enough (H2), a thread per cell;
* ``grid``         workgroups of the persistent step kernel: the resident
  capacity (occupancy x CUs) or a fraction of it (fewer workgroups leave CUs
  idle but reduce contention for the L2 and the work counter at small sizes);
* ``jacobian``     the plain step's Jacobian: difference quotients (the
  reference's CVODE) or analytic -- both reach the same solution; the faster
  one is kept;
* the surrogate-guided solve (when a trained network exists): schedule
  (``surr_adaptive``, ``surr_nsub``, ``surr_grade``, ``surr_max_att``,
  ``surr_h0_fac``), Jacobian updates on slow Newton contraction
  (``surr_max_jac`` 1 or 4 at ``surr_jac_theta`` 0.3),
  network use (``surr_guess``), Newton stop (``surr_newton_tol``) and
  acceptance threshold (``surr_tol_y`` 1.5e-7 .. 2e-6): every combination is
  timed on freshly generated plus relaxed 32^3 states (the block after one
  plain step) and screened: max species error against an FP64 reference --
  the plain step at rtol 1e-10 / atol 1e-16 on the same states -- <= 5e-7
  and max temperature error <= 1e-2 K on both, and at least 1.03x faster
  than the plain step.  The fastest screened ones (up to 24) then run on the
  mechanism's full benchmark block (GRI-3.0 128^3) of the tuner's seed
  against the tight reference there; the first within 6.5e-7 in Y and 1e-2 K
  (larger blocks reach further into the tail, and other seeds' blocks up to
  1.3x further) is recorded.

The state sample is the mechanism's benchmark distribution; for GRI-3.0 the
uneven-stiffness one (70 % near equilibrium, 20 % ignition front, 10 % cold).

Results are cached in ``tiger_nga_amd/tuning.json`` keyed by mechanism, device
name and the sample size bucket; ``apply(fc)`` sets the cached choices on a
FiniteChem.  Usage (on a GPU): ``python -m tiger_nga_amd.autotune --mech gri30``.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")
WORKLOAD = {"h2": "h2", "gri30": "gri30_uneven", "dme39": "dme39", "ethyred": "ethyred"}
# acceptance of a guided-solve configuration against the tight reference:
# screening on the tuning states (2 x 32^3 cells) at half the 1e-6 bound, then
# the fastest screened ones on the mechanism's full benchmark block (the
# largest error grows with the number of cells: 1.3-3.3x from the 32^3 sample
# to 128^3 on GRI-3.0), of the tuner's own seed, at 6.5e-7 (blocks of other
# seeds reach up to 1.3x further)
MAX_SPECIES_ERR = 5e-7
FULL_SPECIES_ERR = 6.5e-7
MAX_T_ERR = 1e-2
VALIDATE_K = 24
# (adaptive, sub-steps, grade, attempt budget) of the guided solve
SURR_SCHEDULES = ((1, 8, 2.0, 48), (1, 4, 1.0, 32), (1, 8, 3.0, 64), (1, 8, 3.0, 96), (0, 4, 1.0, 48),
                  (0, 8, 1.5, 48))
# acceptance thresholds: the adaptive schedule's error against the plain step
# grows with the number of steps it takes (each within the threshold), so the
# mechanisms that take many need the smaller ones
SURR_TOLS = (1.5e-7, 2e-7, 2.5e-7, 5e-7, 1e-6, 2e-6)
# first step of the adaptive schedule from the initial rate (0: the schedule's)
SURR_H0 = (0.0, 16.0)
# adaptive schedule: growth factors in [1, keep) keep the step and its
# factors (1.0: every accepted step takes its new size)
SURR_KEEP = (1.5,)
# Newton stop of the stage solves in acceptance tolerances (surr_newton_tol):
# the stage values' convergence error is systematic and adds up over the
# steps, so a tighter stop with a looser step threshold can be both faster and
# more accurate (adaptive schedule without the network)
SURR_NEWTON_TOL = (0.02, 0.03, 0.05, 0.1)
# Jacobians per cell of the adaptive schedule (updates when Newton contracts
# slower than SURR_JAC_THETA with a Jacobian of an earlier state)
SURR_MAX_JAC = (1, 4)
SURR_JAC_THETA = 0.3
# a guided-solve configuration is recorded only when it beats the plain step
# by this factor on the tuning states (a configuration that rejects nearly
# every cell runs the plain step behind a failed solve and times like it)
SURR_MIN_GAIN = 1.03


def _device_name():
    """Architecture and CU count (the marketing name depends on which ids
    table the runtime finds, and differs under a profiler)."""
    import torch
    p = torch.cuda.get_device_properties(0)
    return "%s/%dCU" % (getattr(p, "gcnArchName", "gpu").split(":")[0], p.multi_processor_count)


def _key(mech, cells):
    return "%s|%s|%d" % (mech, _device_name(), 1 << max(0, int(cells - 1).bit_length()))


def _time(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(mech: str, cells: int = 32 ** 3, seed: int = 7, reps: int = 2, log=print):
    """Measure every candidate on `cells` seeded states; returns the record."""
    import torch
    from . import surrogate
    from .finitechem import FiniteChem
    fc = FiniteChem(mech, 1e-6, 1e-12)
    rec = {"mech": mech, "device": _device_name(), "cells": cells, "seed": seed, "measured": {}}
    # batched right-hand side: lane vs wave per state
    _, Y, T, rho, P, dt = inputs.states(WORKLOAD[mech], n=cells, seed=seed)
    sol = torch.from_numpy(inputs.packed_state(Y, T, rho).T.copy()).cuda()
    times = {}
    for mapping in (1, 0):
        fc.set_option("rhs_mapping", mapping)
        times[mapping] = _time(torch, lambda: fc.compute_rhs(sol, P), reps * 5)
    best_map = min(times, key=times.get)
    rec["rhs_mapping"] = best_map
    rec["measured"]["rhs_ms"] = {("lane" if k else "wave"): v * 1e3 for k, v in times.items()}
    log("  rhs: lane %.3f ms, wave %.3f ms -> %s" % (times[1] * 1e3, times[0] * 1e3, "lane" if best_map else "wave"))
    fc.set_option("rhs_mapping", best_map)
    # persistent grid of the step kernel
    R_d, S0, g, Pg = surrogate.block_states(fc, WORKLOAD[mech], cells, seed, 0)
    S = S0.clone()
    props = torch.cuda.get_device_properties(0)
    n_cu = props.multi_processor_count
    cand = {0: None}
    for frac in (0.5, 0.75):
        cand[max(1, int(n_cu * frac))] = frac

    def step():
        S.copy_(S0)
        fc.source_chemistry(R_d, S, g, dt, Pg, Pg, sync=False)
    gt = {}
    for grid in cand:
        fc.set_option("grid", grid)
        gt[grid] = _time(torch, step, reps)
        log("  step grid %s: %.2f ms" % (grid or "resident", gt[grid] * 1e3))
    best_grid = min(gt, key=gt.get)
    # keep the resident grid unless another is clearly faster (2 %)
    if gt[best_grid] > 0.98 * gt[0]:
        best_grid = 0
    rec["grid"] = best_grid
    rec["measured"]["step_ms"] = {str(k): v * 1e3 for k, v in gt.items()}
    fc.set_option("grid", best_grid)
    # the plain step's mapping: a wavefront per cell (2 or, where supported, 4
    # waves per SIMD) or a thread per cell
    mt = {(0, 2): _time(torch, step, reps)}
    if fc.neq == 10:
        fc.set_option("step_waves", 4)
        mt[(0, 4)] = _time(torch, step, reps)
        fc.set_option("step_waves", 2)
        fc.set_option("step_mapping", 1)
        mt[(1, 2)] = _time(torch, step, reps)
        fc.set_option("step_mapping", 0)
    names = {(0, 2): "wave", (0, 4): "wave, 4 per SIMD", (1, 2): "thread"}
    for k, v in mt.items():
        log("  step mapping %s: %.2f ms" % (names[k], v * 1e3))
    best_step_map, best_waves = min(mt, key=mt.get)
    rec["step_mapping"] = best_step_map
    rec["step_waves"] = best_waves
    rec["measured"]["step_mapping_ms"] = {names[k]: v * 1e3 for k, v in mt.items()}
    fc.set_option("step_mapping", best_step_map)
    fc.set_option("step_waves", best_waves)
    # the plain step's Jacobian: difference quotients or analytic (the thread
    # per cell mapping has difference quotients only)
    jt = {}
    for jac in (0, 1):
        fc.set_option("jacobian", jac)
        jt[jac] = _time(torch, step, reps)
        log("  step jacobian %s: %.2f ms" % ("analytic" if jac else "difference quotients", jt[jac] * 1e3))
    best_jac = 1 if jt[1] < 0.98 * jt[0] and jt[1] < mt[(best_step_map, best_waves)] else 0
    rec["jacobian"] = best_jac
    rec["measured"]["step_jacobian_ms"] = {"difference_quotients": jt[0] * 1e3, "analytic": jt[1] * 1e3}
    fc.set_option("jacobian", best_jac)
    # surrogate-guided solve: schedule, network use and acceptance threshold,
    # measured on freshly generated states and on relaxed ones (the block
    # after one plain step, as in a running simulation)
    path = surrogate.model_path(mech)
    if os.path.exists(path):
        nsp = fc.nsp
        sets = []
        for relax in (0, 1):
            Rb, Sb0, gb, Pb = surrogate.block_states(fc, WORKLOAD[mech], cells, seed, relax)
            Sb = Sb0.clone()

            def pstep(Rb=Rb, Sb=Sb, Sb0=Sb0, gb=gb, Pb=Pb):
                Sb.copy_(Sb0)
                fc.source_chemistry(Rb, Sb, gb, dt, Pb, Pb, sync=False)
            tp = _time(torch, pstep, reps)
            # the FP64 reference of these states: the plain step at tight tolerances
            ft = FiniteChem(mech, 1e-10, 1e-16, 200000)
            Sb.copy_(Sb0)
            ft.source_chemistry(Rb, Sb, gb, dt, Pb, Pb)
            ft.close()
            sets.append((Rb, Sb0, gb, Pb, Sb.clone(), tp))
        rec["measured"]["plain_ms"] = {"generated": sets[0][5] * 1e3, "relaxed": sets[1][5] * 1e3}
        rec["error_reference"] = "plain FP64 step at rtol 1e-10 atol 1e-16 on the same states"
        rec["workload"] = WORKLOAD[mech]
        sur = surrogate.Surrogate.load(path)
        sur.attach(fc)
        fc.set_option("surr_probe", 0)
        st_all = {}
        screened = []
        combos = []
        for (adaptive, nsub, grade, att) in SURR_SCHEDULES:
            for guess in (0, 1):
                # the rate-based first step only for the adaptive schedule
                # without the network (the network's first-step guess assumes
                # the schedule's first sub-step)
                for h0f in (SURR_H0 if adaptive and not guess else (0.0,)):
                    for mj in (SURR_MAX_JAC if adaptive else (1,)):
                        for keep in (SURR_KEEP if adaptive and not guess else (1.5,)):
                            for ntol in (SURR_NEWTON_TOL if adaptive and not guess else (0.1,)):
                                combos.append((adaptive, nsub, grade, att, guess, h0f, mj, keep, ntol))
        fc.set_option("surr_jac_theta", SURR_JAC_THETA)
        for (adaptive, nsub, grade, att, guess, h0f, mj, keep, ntol) in combos:
            for tol in SURR_TOLS:
                fc.set_option("surr_newton_tol", ntol)
                fc.set_option("surr_max_jac", mj)
                fc.set_option("surr_keep", keep)
                fc.set_option("surr_adaptive", adaptive)
                fc.set_option("surr_nsub", nsub)
                fc.set_option("surr_grade", grade)
                fc.set_option("surr_max_att", att)
                fc.set_option("surr_guess", guess)
                fc.set_option("surr_h0_fac", h0f)
                fc.set_option("surr_tol_y", tol)
                t_tot, errs, terrs = 0.0, [], []
                for (Rb, Sb0, gb, Pb, ref, _) in sets:
                    Sb = Sb0.clone()

                    def sstep(Rb=Rb, Sb=Sb, Sb0=Sb0, gb=gb, Pb=Pb):
                        Sb.copy_(Sb0)
                        fc.source_chemistry_surrogate(Rb, Sb, gb, dt, Pb, Pb, sync=False)
                    t_tot += _time(torch, sstep, reps)
                    sstep()
                    torch.cuda.synchronize()
                    errs.append(float((Sb[..., :nsp] - ref[..., :nsp]).abs().max()))
                    terrs.append(float((Sb[..., nsp] - ref[..., nsp]).abs().max()))
                name = surr_name(adaptive, nsub, grade, att, guess, h0f, mj, keep, tol, ntol)
                st_all[name] = {"ms_generated_plus_relaxed": t_tot * 1e3, "max_species_err_generated": errs[0],
                                "max_species_err_relaxed": errs[1], "max_T_err_K": max(terrs)}
                log("  surrogate %s: %.2f ms, max |dY| %.2e / %.2e, max |dT| %.2e K" % (
                    name, t_tot * 1e3, errs[0], errs[1], max(terrs)))
                if max(errs) <= MAX_SPECIES_ERR and max(terrs) <= MAX_T_ERR:
                    screened.append((t_tot, adaptive, nsub, grade, att, guess, h0f, mj, keep, tol, ntol))
        rec["measured"]["surrogate"] = st_all
        plain_s = sets[0][5] + sets[1][5]
        screened = [c for c in sorted(screened) if c[0] * SURR_MIN_GAIN <= plain_s]
        validation = {}
        if screened:
            # the fastest screened candidates on the full benchmark block
            wl = WORKLOAD[mech]
            full_n = inputs.WORKLOADS[wl][1]
            Rf, Sf0, gf, Pf = surrogate.block_states(fc, wl, full_n, seed, 0)
            ft = FiniteChem(mech, 1e-10, 1e-16, 200000)
            ref_f = Sf0.clone()
            ft.source_chemistry(Rf, ref_f, gf, dt, Pf, Pf)
            ft.close()
            for c in screened[:VALIDATE_K]:
                (_, adaptive, nsub, grade, att, guess, h0f, mj, keep, tol, ntol) = c
                for k, v in (("surr_newton_tol", ntol), ("surr_max_jac", mj), ("surr_keep", keep),
                             ("surr_adaptive", adaptive), ("surr_nsub", nsub), ("surr_grade", grade),
                             ("surr_max_att", att), ("surr_guess", guess), ("surr_h0_fac", h0f), ("surr_tol_y", tol)):
                    fc.set_option(k, v)
                Sf = Sf0.clone()
                fc.source_chemistry_surrogate(Rf, Sf, gf, dt, Pf, Pf)
                e = float((Sf[..., :nsp] - ref_f[..., :nsp]).abs().max())
                te = float((Sf[..., nsp] - ref_f[..., nsp]).abs().max())
                name = surr_name(adaptive, nsub, grade, att, guess, h0f, mj, keep, tol, ntol)
                validation[name] = {"max_species_err": e, "max_T_err_K": te,
                                    "passed": e <= FULL_SPECIES_ERR and te <= MAX_T_ERR}
                log("  full block (%d cells) %s: max |dY| %.2e, max |dT| %.2e K" % (full_n, name, e, te))
                del Sf
                if validation[name]["passed"]:
                    break
            del Rf, Sf0, ref_f
            rec["measured"]["validation"] = {"cells": full_n, "seed": seed, "candidates": validation}
        best = guided_choice(screened, validation)
        if not best:
            rec["measured"]["surrogate_choice"] = (
                "none: no configuration within the screening bound is %.2fx faster than the plain step and "
                "within %.1e on the full block" % (SURR_MIN_GAIN, FULL_SPECIES_ERR))
        if best:
            (_, rec["surr_adaptive"], rec["surr_nsub"], rec["surr_grade"], rec["surr_max_att"], rec["surr_guess"],
             rec["surr_h0_fac"], rec["surr_max_jac"], rec["surr_keep"], rec["surr_tol_y"],
             rec["surr_newton_tol"]) = best
            rec["surr_jac_theta"] = SURR_JAC_THETA
            rec["measured"]["surrogate_speedup"] = plain_s / best[0]
        fc.set_surrogate(None, None, None, None)
    fc.close()
    return rec


def guided_choice(screened, validation):
    """The fastest screened candidate (tuples of time and options, fastest
    first, each already within the screening bounds and SURR_MIN_GAIN faster
    than the plain step) that passed on the full block, or None."""
    for c in screened[:VALIDATE_K]:
        v = validation.get(surr_name(*c[1:]))
        if v is not None and v["passed"]:
            return c
    return None


def surr_name(adaptive, nsub, grade, att, guess, h0f, mj, keep, tol, ntol=None):
    """Key of a guided-solve candidate in a record's measurements (records
    made before the Newton stop was swept have no ntol part)."""
    name = "%s n%d r%.1f a%d guess%d h0f%g jac%d keep%.1f tol%.1e" % (
        "adaptive" if adaptive else "fixed", nsub, grade, att, guess, h0f, mj, keep, tol)
    return name if ntol is None else name + " ntol%g" % ntol


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save(rec, cells, path=CACHE):
    cache = load_cache(path)
    cache[_key(rec["mech"], cells)] = rec
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def lookup(mech: str, cells: int, path=CACHE):
    """The cached record for this mechanism and device at the nearest size, or None."""
    cache = load_cache(path)
    dev = _device_name()
    recs = [r for r in cache.values() if r.get("mech") == mech and r.get("device") == dev]
    if not recs:
        return None
    return min(recs, key=lambda r: abs((r["cells"] or 1) - cells))


def apply(fc, cells: int, path=CACHE) -> dict:
    """Set the tuned choices on a FiniteChem; returns what was applied."""
    rec = lookup(fc.mech, cells, path)
    done = {}
    if not rec:
        return done
    for k in ("rhs_mapping", "grid", "step_mapping", "step_waves", "jacobian", "surr_adaptive", "surr_nsub", "surr_grade", "surr_max_att",
              "surr_guess", "surr_h0_fac", "surr_jac_theta", "surr_max_jac", "surr_keep", "surr_tol_y",
              "surr_newton_tol"):
        if k in rec:
            fc.set_option(k, rec[k])
            done[k] = rec[k]
    return done


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", nargs="*", default=["h2", "gri30", "dme39", "ethyred"])
    ap.add_argument("--cells", type=int, default=32 ** 3)
    ap.add_argument("--cache", default=CACHE)
    a = ap.parse_args(argv)
    for m in a.mech:
        print("== %s" % m, flush=True)
        rec = tune(m, a.cells, log=lambda s: print(s, flush=True))
        save(rec, a.cells, a.cache)
        print(json.dumps({k: v for k, v in rec.items() if k != "measured"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
