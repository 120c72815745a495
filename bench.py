#!/usr/bin/env python3
"""Chemistry-step benchmark on one MI355X.

Workload (default ``gri30``): GRI-Mech 3.0 as shipped with the reference
(grimech30F, 36 species, 219 reactions / 422 rate slots) on 64^3 = 262,144
cells at 1 atm, dt = 5e-7 s, seeded synthetic states (tiger_nga_amd/inputs.py),
plain FP64 path at the reference tolerances (rtol 1e-6, atol 1e-12,
MAX_NSTEPS 10000).  One step = one finitechem_source_chemistry call over all
cells with the fields resident in HBM (device entry point); before each step
the fields are restored from a pristine device copy so that every step does
the same work.

Prints one JSON line: cell-steps/s, the FP64 roofline of the step kernel
(algorithmic flops from the integrator counters / kernel time, HIP events on
the launch stream) and the CPU baseline: the C oracle (CVODE-semantics BDF and
table-driven mechanism) on a bounded sample of the same cells, one cell loop
per host thread.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="gri30", choices=["h2", "gri30", "gri30_uneven", "dme39", "ethyred"])
    ap.add_argument("--cpu-cells", type=int, default=131072)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all usable host cores, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default="step", choices=["step", "properties"],
                    help="step: the chemistry step (the headline); properties: density / viscosity / "
                         "diffusivities of the same block")
    ap.add_argument("--data-file", default=None,
                    help="step the fields of an NGA data file instead of the synthetic workload "
                         "(needs --mech; no CPU baseline)")
    ap.add_argument("--mech", default=None, choices=["h2", "gri30", "dme39", "ethyred"], help="mechanism of --data-file")
    ap.add_argument("--pressure", type=float, default=101325.0, help="P_thermo for --data-file (Pa)")
    ap.add_argument("--dt", type=float, default=0.0, help="chemistry dt for --data-file (0: the file's dt)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate-guided step (timed when a trained network for the mechanism exists)")
    return ap.parse_args()


def fortran_empty(torch, shape, dev):
    strides, acc = [], 1
    for s in shape:
        strides.append(acc)
        acc *= s
    return torch.empty_strided(shape, tuple(strides), dtype=torch.float64, device=dev)


def cpu_baseline(workload, ncell, nthreads):
    import numpy as np

    from oracle import pyoracle
    from tiger_nga_amd import inputs
    # the benchmark's own states (the generator draws per field over all
    # cells, so a shorter run of it would give other states): its first cells
    mech, Y, T, rho, P, dt = inputs.states(workload, seed=0)
    ncell = min(ncell, Y.shape[0])
    sol = inputs.packed_state(Y[:ncell], T[:ncell], rho[:ncell])
    t0 = time.perf_counter()
    result = pyoracle.integrate_batch(mech, sol, dt, P, rtol=1e-6, atol=1e-12, nthreads=nthreads)
    el = time.perf_counter() - t0
    n1 = max(16, ncell // 64)
    t0 = time.perf_counter()
    pyoracle.integrate_batch(mech, sol[:n1], dt, P, rtol=1e-6, atol=1e-12, nthreads=1)
    el1 = time.perf_counter() - t0
    rec = {"value": ncell / el, "unit": "cell-steps/s", "cores": nthreads, "kind": "port",
           "sample": "first %d cells of the %s workload (same seeded states), C oracle BDF at rtol 1e-6, "
                     "%d threads x serial cell loop; %.2f s wall" % (ncell, workload, nthreads, el),
           "single_core_value": n1 / el1}
    # oracle result as mass fractions and T, for the error of the GPU paths
    rho = result[:, :-1].sum(axis=1)
    return rec, np.concatenate([result[:, :-1] / rho[:, None], (result[:, -1] / rho)[:, None]], axis=1)


def errors_vs_oracle(torch, S, nsp, ref):
    """Max / RMS mass-fraction error and max T error of the first len(ref)
    interior cells of a Fortran-strided (nx, ny, nz, nscalar) result."""
    n = ref.shape[0]
    got = torch.stack([S[..., k].permute(2, 1, 0).reshape(-1)[:n] for k in range(nsp + 1)], dim=1).cpu().numpy()
    dY = abs(got[:, :nsp] - ref[:, :nsp])
    return {"cells": n, "max_species_err": float(dY.max()), "rms_species_err": float((dY ** 2).mean() ** 0.5),
            "max_T_err_K": float(abs(got[:, nsp] - ref[:, nsp]).max())}


def traffic_record(workload, kernel="tgr::step_kernel"):
    """HBM bytes per launch of a kernel (name prefix) for this workload, from
    the committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE runs,
    scaled by a copy kernel of known size); None when no summary covers it."""
    import glob
    for f in sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "*_traffic.json"))):
        with open(f) as fh:
            rec = json.load(fh)
        if rec.get("workload") == workload and rec.get("kernel", "").startswith(kernel):
            return rec
    return None


def _progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print("[bench %s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)


def _time_steps(torch, a, fn, S_d, S0_d):
    """Seconds per step of fn over K steps (state restored before each)."""
    for _ in range(max(1, a.warmup)):
        S_d.copy_(S0_d)
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        S_d.copy_(S0_d)
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / a.steps


def bench_surrogate(a, torch, fc, g, R_d, S0_d, S_d, HR_d, CS_d, dt, P, mech, sp, stream, cells):
    """The surrogate-guided step on the same block (S_d holds the plain FP64
    result on entry): forced guided solve on every call, the adaptive policy
    (guided solve only while it pays, with periodic probes), the same guided
    solve without the network (Newton iterations compared), and both paths on
    relaxed states (the block after one plain step, as in a running
    simulation)."""
    from tiger_nga_amd import surrogate
    path = surrogate.model_path(mech)
    if a.no_surrogate or not os.path.exists(path):
        return None
    nsp = fc.nsp
    Sp = S_d.clone()
    sur = surrogate.Surrogate.load(path)
    sur.attach(fc)

    def guided():
        return fc.source_chemistry_surrogate(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d, stream=sp, sync=False)

    def errors(ref):
        dY = (S_d[..., :nsp] - ref[..., :nsp]).abs()
        dT = (S_d[..., nsp] - ref[..., nsp]).abs()
        return float(dY.max()), float(dY.pow(2).mean().sqrt()), float(dT.max())

    from tiger_nga_amd import flops
    guess_t = int(fc.options.get("surr_guess", 0))  # the tuned (or default) network use
    fc.set_option("surr_probe", 0)
    _progress("guided step, forced on every call")
    t_forced = _time_steps(torch, a, guided, S_d, S0_d)
    S_d.copy_(S0_d)
    _, ss = fc.source_chemistry_surrogate(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d, stream=sp)
    e_max, e_rms, eT = errors(Sp)
    state = S_d.clone()  # the guided result on the benchmark states, for the oracle comparison
    # the other network setting on the same states: what the network's
    # guesses gain (or cost)
    fc.set_option("surr_guess", 1 - guess_t)
    _progress("guided step with surr_guess %d" % (1 - guess_t))
    t_other = _time_steps(torch, a, guided, S_d, S0_d)
    S_d.copy_(S0_d)
    _, so = fc.source_chemistry_surrogate(R_d, S_d, g, dt, P, P, stream=sp)
    o_max = errors(Sp)[0]
    fc.set_option("surr_guess", guess_t)
    fc.set_option("surr_probe", 8)
    _progress("guided step under the adaptive-use policy")
    t_adapt = _time_steps(torch, a, guided, S_d, S0_d)
    # relaxed states: both paths
    R2 = S0_d.clone()
    S_d.copy_(S0_d)
    fc.source_chemistry(R_d, S_d, g, dt, P, P, stream=sp)
    R2.copy_(S_d)
    _progress("relaxed states: plain and guided steps")
    t_rp = _time_steps(torch, a, lambda: fc.source_chemistry(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d,
                                                            stream=sp, sync=False), S_d, R2)
    Rp = S_d.clone()
    fc.set_option("surr_probe", 0)
    t_rg = _time_steps(torch, a, guided, S_d, R2)
    S_d.copy_(R2)
    _, sr = fc.source_chemistry_surrogate(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d, stream=sp)
    r_max, r_rms, rT = errors(Rp)
    fc.set_option("surr_probe", 8)
    fc.set_surrogate(None, None, None, None)
    gf = (flops.guided_flops(fc.info, ss, split=bool(fc.options.get("surr_split_jac", 1)))
          if ss["surrogate_ms"] > 0 else 0.0)

    def net_case(t, st, emax):
        return {"value": cells / t, "ms_per_step": t * 1e3, "accepted_fraction": st["accepted"] / max(1, cells),
                "newton_iters_per_cell": st["newton_iters"] / max(1, cells), "max_species_err": emax}
    cur, oth = net_case(t_forced, ss, e_max), net_case(t_other, so, o_max)
    with_net, without_net = (cur, oth) if guess_t else (oth, cur)
    opts = {k: fc.options[k] for k in sorted(fc.options) if k.startswith("surr_")}
    return {
        "_state": state,
        "guided_roofline": {"bound": "mfma", "compute_unit": "FP64 vector ALU", "achieved": gf / (
            ss["surrogate_ms"] * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": gf / (ss["surrogate_ms"] * 1e-3) / PEAK_FP64, "flops": gf, "kernel_ms": ss["surrogate_ms"],
            "kernel": "tgr::surr_kernel",
            "note": "HIP events on the launch stream around the guided part (surr_kernel; with the network "
                    "guess also the feature and network launches); flops of the FP64 guided solve"},
        "value": cells / t_adapt, "ms_per_step": t_adapt * 1e3, "model": os.path.relpath(path, ROOT),
        "policy": "adaptive: guided solve while its accepted fraction >= 0.15, else plain step with a guided "
                  "probe every 8 calls",
        "forced_guided": {"value": cells / t_forced, "ms_per_step": t_forced * 1e3,
                          "accepted_fraction": ss["accepted"] / max(1, cells),
                          "rejected": {"newton": ss["rej_newton"], "error_estimate": ss["rej_error"],
                                       "enthalpy": ss["rej_enthalpy"]},
                          "guided_ms": ss["surrogate_ms"], "fallback_ms": ss["fallback_ms"],
                          "newton_iters_per_cell": ss["newton_iters"] / max(1, cells),
                          "substeps_per_cell": ss["substeps"] / max(1, cells)},
        "options": opts,
        "network": {"used": bool(guess_t), "with_network_guess": with_net, "without_network_guess": without_net,
                    "gain": without_net["ms_per_step"] / with_net["ms_per_step"]},
        "max_species_err": e_max, "rms_species_err": e_rms, "max_T_err_K": eT,
        "error_reference": "plain FP64 step of this run on the same states (rtol 1e-6, atol 1e-12)",
        "relaxed_states": {"what": "the same block after one plain step of dt (radical pools relaxed)",
                           "plain_value": cells / t_rp, "guided_value": cells / t_rg, "speedup": t_rp / t_rg,
                           "accepted_fraction": sr["accepted"] / max(1, cells), "max_species_err": r_max,
                           "rms_species_err": r_rms, "max_T_err_K": rT},
        "method": "per cell: 3-stage L-stable SDIRK3(2) with one frozen FP64 Jacobian, %s sub-steps, FP64 "
                  "Newton from %s, embedded error <= surr_tol_y in Y, element-conserving projection, "
                  "enthalpy-consistent T; rejected cells by the plain FP64 step"
                  % ("error-controlled" if fc.options.get("surr_adaptive", 1) else "fixed",
                     "the network's guesses" if guess_t else "the derivative predictor")}


def bench_properties(a, torch, fc, g, S_d, R_d, P, n, cells_label, mech, sp):
    """Density, viscosity and diffusivities of every cell (finitechem.f90:324-461)."""
    from tiger_nga_amd import flops
    V_d = torch.empty_like(R_d)
    D_d = torch.empty_like(S_d)
    for _ in range(max(1, a.warmup)):
        fc.properties(S_d, g, P, RHO_=R_d, VISC_=V_d, DIFF_=D_d, stream=sp)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        fc.properties(S_d, g, P, RHO_=R_d, VISC_=V_d, DIFF_=D_d, stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ks = e0.elapsed_time(e1) / 1e3 / a.steps
    F = flops.properties_flops(fc.nsp) * n
    B = 8.0 * n * ((fc.nsp + 1) + (fc.nsp + 3))
    out = {"metric": "transport/EOS property cell-updates/s, 1 GPU", "value": n * a.steps / el, "unit": "cells/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": el / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "%s density + Wilke viscosity + diffusivities, %s" % (mech, cells_label)},
           "roofline": {"bound": "mfma", "compute_unit": "FP64 vector ALU", "achieved": F / ks / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": F / ks / PEAK_FP64, "traffic": None,
                        "kernel_ms": ks * 1e3, "flops_per_step": F, "algorithmic_bytes_per_step": B,
                        "hbm_frac_at_algorithmic_bytes": B / ks / PEAK_HBM}}
    print(json.dumps(out), flush=True)
    return 0


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the chemistry step has no cross-GPU work in this build "
              "(cells are independent); run with --gpus 1", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from tiger_nga_amd import flops, inputs
    from tiger_nga_amd.finitechem import FiniteChem, Grid

    dev = "cuda:0"
    torch.cuda.set_device(0)
    if a.data_file:
        from tiger_nga_amd import ngadata
        from tiger_nga_amd.mech import meta
        if not a.mech:
            print("--data-file needs --mech", file=sys.stderr)
            return 2
        df = ngadata.read(a.data_file)
        mech, P = a.mech, a.pressure
        dt = a.dt or df.dt
        if not dt > 0.0:
            print("the data file holds dt = %g: give --dt" % df.dt, file=sys.stderr)
            return 2
        R_d, S0_d, g = ngadata.chemistry_block(df, meta.species(mech), device=dev)
        n = df.nx * df.ny * df.nz
        cells_label = "%dx%dx%d = %d cells" % (df.nx, df.ny, df.nz, n)
        workload = "data file %s" % os.path.basename(a.data_file)
        a.no_cpu = True
    else:
        mech, Y, T, rho, P, dt = inputs.states(a.workload, seed=0)
        n = Y.shape[0]
        nx = int(round(n ** (1.0 / 3.0)))
        assert nx ** 3 == n
        RHO_, SC_, isc_sc, isc_T = inputs.fortran_fields(Y, T, rho, nx, nx, nx)
        g = Grid.block(nx, nx, nx, SC_.shape[3], isc_sc, isc_T)
        R_d = fortran_empty(torch, RHO_.shape, dev)
        S0_d = fortran_empty(torch, SC_.shape, dev)
        R_d.copy_(torch.from_numpy(RHO_))
        S0_d.copy_(torch.from_numpy(SC_))
        del RHO_, SC_, Y, T, rho
        cells_label = "%d^3 = %d cells" % (nx, n)
        workload = a.workload
    S_d = fortran_empty(torch, tuple(S0_d.shape), dev)
    HR_d = fortran_empty(torch, tuple(R_d.shape), dev)
    CS_d = fortran_empty(torch, tuple(R_d.shape) + (3,), dev)
    fc = FiniteChem(mech, 1e-6, 1e-12, 10000)
    from tiger_nga_amd import autotune
    tuned = autotune.apply(fc, n)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    if a.kernel == "properties":
        return bench_properties(a, torch, fc, g, S0_d, R_d, P, n, cells_label, mech, sp)

    stats = None
    _progress("plain step, %s" % cells_label)
    for _ in range(max(1, a.warmup)):
        S_d.copy_(S0_d)
        stats = fc.source_chemistry(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d, stream=sp, sync=True)
    torch.cuda.synchronize()

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        S_d.copy_(S0_d)
        evs[k][0].record(stream)
        fc.source_chemistry(R_d, S_d, g, dt, P, P, HR_gp=HR_d, chemsrc_gp=CS_d, stream=sp, sync=False)
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kms = [e0.elapsed_time(e1) for e0, e1 in evs]
    kernel_s = sum(kms) / len(kms) / 1e3

    info = fc.info
    F = flops.step_flops(info, stats)
    achieved = F / kernel_s
    cells = stats["cells"]
    metric = "GRI-3.0 chem cell-steps/s, 1 GPU: surrogate-guided vs plain FP64 at err<=1e-6"
    if mech != "gri30":
        metric = metric.replace("GRI-3.0", {"h2": "HYDROGENF", "dme39": "DME39F", "ethyred": "EthyRedF"}[mech])
    out = {
        "metric": metric,
        "value": cells * a.steps / el,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": ("fields of %s" % a.data_file) if a.data_file else
                "synthetic seeded states (no case data ships with the reference)",
        "config": {"workload": "%s plain FP64 step, %s, dt %.1e s, P %.0f Pa, rtol 1e-6 atol 1e-12"
                               % (workload, cells_label, dt, P), "mechanism": mech, "path": "plain_fp64",
                   "species": info["nsp"], "reactions": info["nrxn"], "rate_slots": info["nslot"],
                   "tuned": tuned or "defaults (no tuning record for this device)"},
        "roofline": {"bound": "mfma", "compute_unit": "FP64 vector ALU (MI355X FP64 vector peak = FP64 matrix "
                                                       "peak = 78.6 TFLOP/s)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": None,
                     "kernel_ms": kernel_s * 1e3, "flops_per_step": F,
                     "algorithmic_bytes_per_step": flops.step_bytes(info, cells)},
        "solver": {k: stats[k] for k in ("nst", "nfe", "nfe_dq", "nje", "nsetups", "nni", "netf", "ncfn", "nlu",
                                         "failed")},
    }
    S_plain = S_d.clone()
    sur = bench_surrogate(a, torch, fc, g, R_d, S0_d, S_d, HR_d, CS_d, dt, P, mech, sp, stream, cells)
    sur_state = None
    if sur is not None:
        sur_state = sur.pop("_state")
        out["surrogate"] = sur
        # the metric compares the surrogate-guided step with the plain FP64 one
        # at max species error <= 1e-6: the guided path is the headline when
        # it meets that bound, else the plain path
    tr = None if a.data_file else traffic_record(a.workload)
    if tr:
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        out["roofline"]["traffic_source"] = tr["source"]
    if sur is not None and sur["max_species_err"] <= 1e-6:
        # the guided path is the headline: its throughput, and the roofline
        # of its dominant kernel (the guided solve); the plain step's beside
        out["plain_value"] = out["value"]
        out["value"] = sur["value"]
        out["ms_per_step"] = sur["ms_per_step"]
        out["config"]["path"] = "surrogate_guided"
        out["config"]["workload"] = ("%s chemistry step, SDIRK3(2)-guided with the plain FP64 step on rejected cells, "
                                     "%s, dt %.1e s, P %.0f Pa, error <= 1e-6 against rtol 1e-6 atol 1e-12"
                                     % (workload, cells_label, dt, P))
        out["plain_roofline"] = out["roofline"]
        gr = dict(sur["guided_roofline"])
        gr["traffic"] = None
        trs = None if a.data_file else traffic_record(a.workload, "tgr::surr_kernel")
        if trs:
            gr["traffic"] = trs["hbm_bytes_per_launch"]
            gr["traffic_source"] = trs["source"]
        out["roofline"] = gr
    if not a.no_cpu:
        nt = a.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        _progress("CPU oracle on %d cells, %d threads" % (a.cpu_cells, nt))
        out["cpu_baseline"], ref = cpu_baseline(a.workload, a.cpu_cells, nt)
        # both GPU paths against the FP64 oracle on the oracle's cells
        out["plain_vs_oracle"] = errors_vs_oracle(torch, S_plain, fc.nsp, ref)
        if sur is not None:
            out["surrogate"]["vs_oracle"] = errors_vs_oracle(torch, sur_state, fc.nsp, ref)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
