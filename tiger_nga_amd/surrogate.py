"""The learned part of the chemistry step: training data from the plain FP64
step, training of the BF16-MFMA perceptron on the GPU, and the model file.

Data: the plain device step (the FP64 solver that is also the fallback)
advances seeded synthetic states (tiger_nga_amd/inputs.py; seeds disjoint
from the benchmark's seed 0) by several time steps; each cell gives one pair
(network input of the initial state and dt) -> (increment of Y_k and T over
dt).  Inputs are computed by the same device kernel that feeds the network in
the surrogate step (tgr_chem_surrogate_features_device), so training and use
see identical features.  Targets are increments divided by out_scale, the
per-component RMS increment of the training set.

Model file (``.bin``, little-endian numbers only, read by the C ABI's
tgr_chem_load_surrogate as well): the 8 bytes ``TGRMLP01``; int32 nin, nout,
hidden, layers; int64 nparam; float32 in_shift[nin], in_scale[nin]; float64
out_scale[nout]; float32 params[nparam] (the flat vector of
include/tiger_mlp.h).  The training report sits beside it as ``.json``.

Usage (on a GPU): ``python -m tiger_nga_amd.surrogate --mech gri30 --out
tiger_nga_amd/models/gri30.bin``.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import inputs
from .mech import meta

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "models")
# time steps of the training pairs as fractions of the workload's dt: a
# geometric ladder over the stage times of graded guided solves (down to
# gamma * dt / 63 for six sub-steps of ratio 2)
DT_FRACTIONS = tuple(float(v) for v in np.geomspace(0.005, 1.25, 8))
Y_FLOOR = 1.0e-30

TRAIN_SETS = {
    # mechanism: [(workload of the state generator, cells per launch, relax)]
    # relax = 1: the generated states are first advanced by one plain step of
    # the workload's dt (radical pools relaxed, as in a running simulation)
    "h2": [("h2", 64 ** 3, 0), ("h2", 64 ** 3, 1)],
    "gri30": [("gri30", 64 ** 3, 0), ("gri30_uneven", 64 ** 3, 0), ("gri30", 64 ** 3, 1),
              ("gri30_uneven", 64 ** 3, 1)],
    "dme39": [("dme39", 64 ** 3, 0), ("dme39", 64 ** 3, 1)],
    "ethyred": [("ethyred", 64 ** 3, 0), ("ethyred", 64 ** 3, 1)],
}


MAGIC = b"TGRMLP01"


def model_path(mech: str) -> str:
    return os.path.join(MODELS, "%s.bin" % mech)


def input_scaling(mech: str, Y, T, dt_ref: float):
    """Shift / scale of the inputs: mean and 1/std of log10 max(Y_k, 1e-30), T,
    and of log10 dt over the training fractions of dt_ref."""
    nsp = len(meta.species(mech))
    ly = np.log10(np.maximum(Y, Y_FLOOR))
    shift = np.zeros(nsp + 2)
    scale = np.ones(nsp + 2)
    shift[:nsp] = ly.mean(axis=0)
    scale[:nsp] = 1.0 / np.maximum(ly.std(axis=0), 1e-3)
    shift[nsp] = T.mean()
    scale[nsp] = 1.0 / max(T.std(), 1e-3)
    ldt = np.log10(np.array(DT_FRACTIONS) * dt_ref)
    shift[nsp + 1] = ldt.mean()
    scale[nsp + 1] = 1.0 / max(ldt.std(), 1e-3)
    return shift.astype(np.float32), scale.astype(np.float32)


def _fortran_dev(torch, a, dev):
    st, acc = [], 1
    for s in a.shape:
        st.append(acc)
        acc *= s
    t = torch.empty_strided(a.shape, tuple(st), dtype=torch.float64, device=dev)
    t.copy_(torch.from_numpy(a))
    return t


def _interior_rows(torch, S, nvar):
    """[cells, nvar] from a Fortran-strided (nx, ny, nz, nvar) tensor, i fastest."""
    return torch.stack([S[..., k].permute(2, 1, 0).reshape(-1) for k in range(nvar)], dim=1)


def block_states(fc, workload: str, n: int, seed: int, relax: int = 0):
    """(RHO_, SC_, Grid, P) of a cubic block of seeded states on the GPU;
    relax = 1 advances them first by one plain step of the workload's dt."""
    import torch
    from .finitechem import Grid
    _, Y, T, rho, P, dt = inputs.states(workload, n=n, seed=seed)
    nx = round(n ** (1.0 / 3.0))
    RHO_, SC_, isc_sc, isc_T = inputs.fortran_fields(Y, T, rho, nx, nx, nx)
    g = Grid.block(nx, nx, nx, SC_.shape[3], isc_sc, isc_T)
    R_d, S_d = _fortran_dev(torch, RHO_, "cuda"), _fortran_dev(torch, SC_, "cuda")
    if relax:
        fc.source_chemistry(R_d, S_d, g, dt, P, P)
    return R_d, S_d, g, P


def generate(fc, mech: str, seeds, dt_ref: float, in_shift, in_scale, launches=None, log=print):
    """(X bf16 [N, in_pad], D float64 [N, neq] increments of Y and T) on the GPU."""
    import torch
    from .finitechem import Grid
    from .mlp import MLP
    nsp = fc.nsp
    # a stand-in network so that the feature kernel has the scaling (the
    # features do not depend on the network's parameters)
    probe = MLP(nsp + 2, nsp + 1, 64, 1)
    fc.set_surrogate(probe, in_shift, in_scale, np.ones(nsp + 1))
    Xs, Ds = [], []
    sets = TRAIN_SETS[mech]
    # one launch per (seed, state set), the time steps cycling through DT_FRACTIONS
    plan = launches or [(wl, n, relax, seed, DT_FRACTIONS[(i * len(sets) + w) % len(DT_FRACTIONS)])
                        for i, seed in enumerate(seeds) for w, (wl, n, relax) in enumerate(sets)]
    for (wl, n, relax, seed, frac) in plan:
        R_d, S_d, g, P = block_states(fc, wl, n, seed, relax)
        rows = (n + 63) // 64 * 64
        X = torch.empty((rows, probe.in_pad), dtype=torch.bfloat16, device="cuda")
        dt = frac * dt_ref
        fc.surrogate_features(S_d, g, dt, X)
        before = _interior_rows(torch, S_d, nsp + 1)
        t0 = time.perf_counter()
        st = fc.source_chemistry(R_d, S_d, g, dt, P, P)
        after = _interior_rows(torch, S_d, nsp + 1)
        ok = torch.ones(n, dtype=torch.bool, device="cuda")
        if st["failed"]:
            ok &= torch.isfinite(after).all(dim=1)
        Xs.append(X[:n][ok])
        Ds.append((after - before)[ok])
        log("  data %-13s relax %d seed %3d dt %.3e: %d cells, %.2f s, failed %d"
            % (wl, relax, seed, dt, n, time.perf_counter() - t0, st["failed"]))
    fc.set_surrogate(None, None, None, None)
    return torch.cat(Xs), torch.cat(Ds)


def train(net, X, Tg, steps: int, batch: int = 4096, lr: float = 2e-3, lr_end: float = 2e-5, seed: int = 0,
          log=print, log_every: int = 1000):
    """Adam with cosine learning-rate decay on random mini-batches."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    n = X.shape[0]
    t0 = time.perf_counter()
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), device="cuda", generator=gen)
        xb, tb = X[idx].contiguous(), Tg[idx].contiguous()
        rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * it / max(1, steps - 1)))
        want = (it % log_every == 0) or it == steps - 1
        loss = net.train_step(xb, tb, rate, want_loss=want)
        if want:
            hist.append((it, loss))
            log("  step %6d  lr %.2e  batch loss %.4e  (%.1f s)" % (it, rate, loss, time.perf_counter() - t0))
    return hist


def evaluate(net, X, Tg, batch: int = 65536):
    """Mean-squared error of the scaled increments and per-component RMS."""
    import torch
    se = torch.zeros(Tg.shape[1], dtype=torch.float64, device="cuda")
    n = X.shape[0] // 64 * 64
    for r0 in range(0, n, batch):
        r1 = min(n, r0 + batch)
        out = net.forward(X[r0:r1].contiguous())[:, :Tg.shape[1]]
        se += ((out.double() - Tg[r0:r1].double()) ** 2).sum(dim=0)
    mse = se / n
    return float(mse.mean()), mse.sqrt().cpu().numpy()


class Surrogate:
    """A trained network with its scaling, attachable to a FiniteChem."""

    def __init__(self, net, in_shift, in_scale, out_scale, meta_info):
        self.net, self.in_shift, self.in_scale, self.out_scale, self.meta = net, in_shift, in_scale, out_scale, \
            meta_info

    def attach(self, fc):
        fc.set_surrogate(self.net, self.in_shift, self.in_scale, self.out_scale)

    def save(self, path):
        n = self.net
        params = n.get_params()
        with open(path, "wb") as f:
            f.write(MAGIC)
            f.write(np.array([n.nin, n.nout, n.hidden, n.layers], dtype="<i4").tobytes())
            f.write(np.array([params.size], dtype="<i8").tobytes())
            f.write(np.asarray(self.in_shift, dtype="<f4").tobytes())
            f.write(np.asarray(self.in_scale, dtype="<f4").tobytes())
            f.write(np.asarray(self.out_scale, dtype="<f8").tobytes())
            f.write(params.astype("<f4").tobytes())
        with open(os.path.splitext(path)[0] + ".json", "w") as f:
            json.dump(self.meta, f, indent=1)

    @staticmethod
    def read(path):
        """(shape, in_shift, in_scale, out_scale, params, meta) of a model file."""
        with open(path, "rb") as f:
            raw = f.read()
        if raw[:8] != MAGIC:
            raise ValueError("%s: not a network file" % path)
        nin, nout, hidden, layers = (int(v) for v in np.frombuffer(raw, "<i4", 4, 8))
        nparam = int(np.frombuffer(raw, "<i8", 1, 24)[0])
        o = 32
        in_shift = np.frombuffer(raw, "<f4", nin, o)
        in_scale = np.frombuffer(raw, "<f4", nin, o + 4 * nin)
        o += 8 * nin
        out_scale = np.frombuffer(raw, "<f8", nout, o)
        o += 8 * nout
        params = np.frombuffer(raw, "<f4", nparam, o)
        if o + 4 * nparam != len(raw):
            raise ValueError("%s: size does not match its header" % path)
        mp = os.path.splitext(path)[0] + ".json"
        meta_info = json.load(open(mp)) if os.path.exists(mp) else {}
        return (nin, nout, hidden, layers), in_shift.copy(), in_scale.copy(), out_scale.copy(), params.copy(), meta_info

    @staticmethod
    def load(path):
        from .mlp import MLP
        (nin, nout, hidden, layers), in_shift, in_scale, out_scale, params, meta_info = Surrogate.read(path)
        net = MLP(nin, nout, hidden, layers)
        net.set_params(params)
        return Surrogate(net, in_shift, in_scale, out_scale, meta_info)


def build(mech: str, seeds=range(1, 9), steps: int = 20000, hidden: int = 256, layers: int = 4, log=print):
    """Generate data, train, and return (Surrogate, report)."""
    import torch
    from .finitechem import FiniteChem
    from .mlp import MLP, init_params
    wl0 = TRAIN_SETS[mech][0][0]
    dt_ref = inputs.WORKLOADS[wl0][3]
    nsp = len(meta.species(mech))
    # input scaling from the states of the first training seed
    ys, ts = [], []
    for wl, n, _ in TRAIN_SETS[mech]:
        _, Y, T, _, _, _ = inputs.states(wl, n=min(n, 65536), seed=seeds[0])
        ys.append(Y)
        ts.append(T)
    in_shift, in_scale = input_scaling(mech, np.concatenate(ys), np.concatenate(ts), dt_ref)
    fc = FiniteChem(mech, 1e-6, 1e-12)
    t0 = time.perf_counter()
    X, D = generate(fc, mech, list(seeds), dt_ref, in_shift, in_scale, log=log)
    t_data = time.perf_counter() - t0
    out_scale = D.pow(2).mean(dim=0).sqrt().clamp_min(1e-300).cpu().numpy()
    Tg = (D / torch.from_numpy(out_scale).cuda()).float().contiguous()
    # hold out every 16th pair for validation
    n = X.shape[0]
    perm = torch.randperm(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(1))
    nval = n // 16 // 64 * 64
    val, tr = perm[:nval], perm[nval:]
    Xtr, Ttr, Xva, Tva = X[tr].contiguous(), Tg[tr].contiguous(), X[val].contiguous(), Tg[val].contiguous()
    log("training pairs %d, validation %d (data %.1f s)" % (Xtr.shape[0], Xva.shape[0], t_data))
    net = MLP(nsp + 2, nsp + 1, hidden, layers)
    net.set_params(init_params(nsp + 2, nsp + 1, hidden, layers, seed=0))
    t1 = time.perf_counter()
    hist = train(net, Xtr, Ttr, steps, log=log)
    t_train = time.perf_counter() - t1
    mse, rms = evaluate(net, Xva, Tva)
    report = {"mech": mech, "pairs": int(Xtr.shape[0]), "validation_pairs": int(Xva.shape[0]),
              "dt_ref": dt_ref, "dt_fractions": list(DT_FRACTIONS), "seeds": list(seeds), "steps": steps,
              "batch": 4096, "data_s": t_data, "train_s": t_train, "val_mse_scaled": mse,
              "val_rms_scaled_max": float(rms.max()), "loss_history": hist}
    # inert species (N2, Ar) have increments at the rounding level of
    # rho Y / rho: their scaled targets are noise, so the worst component
    # over the reacting species and T is reported beside the overall one
    active = out_scale > 1e-12 * out_scale.max()
    report["val_rms_scaled_max_active"] = float(rms[active].max())
    log("validation MSE (scaled) %.4e, worst component RMS %.3f (%.3f over the reacting species and T)"
        % (mse, rms.max(), report["val_rms_scaled_max_active"]))
    return Surrogate(net, in_shift, in_scale, out_scale, report), report


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="gri30", choices=sorted(TRAIN_SETS))
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--seeds", type=int, default=8, help="seeds 1..N of the state generator")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "training runs on the GPU"
    sur, rep = build(a.mech, seeds=range(1, a.seeds + 1), steps=a.steps, log=lambda s: print(s, flush=True))
    out = a.out or model_path(a.mech)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    sur.save(out)
    print(json.dumps({k: v for k, v in rep.items() if k != "loss_history"}), flush=True)
    print("saved", out, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
