"""The BF16-MFMA perceptron kernels against the FP64 CPU reference
(tests/mlp_ref.py): the product kernel in its four operand layouts, the
forward pass, the loss gradient of every parameter, and Adam training."""
import numpy as np
import pytest

from . import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    return torch


def _bf(torch, a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("a_kcontig,b_kcontig,M,N,K", [
    (True, True, 128, 192, 96), (True, True, 256, 256, 64), (True, True, 64, 128, 32), (True, False, 128, 192, 96),
    (False, True, 128, 192, 96), (False, False, 128, 192, 96)])
def test_gemm_layouts(torch_dev, a_kcontig, b_kcontig, M, N, K):
    """The product kernel in each operand layout and at several tile counts."""
    torch = torch_dev
    from tiger_nga_amd.mlp import MLP
    net = MLP(8, 8, 64, 1)
    rng = np.random.default_rng(7)
    A = rng.standard_normal((M, K))
    Bm = rng.standard_normal((K, N))
    Aq, Bq = mlp_ref.bf16(A), mlp_ref.bf16(Bm)
    if a_kcontig:
        Ad, sam, sak = _bf(torch, A), K, 1
    else:
        Ad, sam, sak = _bf(torch, A.T.copy()), 1, M
    if b_kcontig:
        Bd, sbk, sbn = _bf(torch, Bm.T.copy()), 1, K
    else:
        Bd, sbk, sbn = _bf(torch, Bm), N, 1
    C = torch.full((M, N), np.nan, dtype=torch.float32, device="cuda")
    net.gemm(Ad, sam, sak, Bd, sbk, sbn, M, N, K, C, N)
    torch.cuda.synchronize()
    ref = Aq @ Bq
    scale = np.abs(Aq) @ np.abs(Bq)
    err = np.abs(C.cpu().numpy() - ref) / scale
    assert err.max() < 1e-6, err.max()


def _net_case(torch, nin=38, nout=37, hidden=256, layers=4, B=128, seed=5):
    from tiger_nga_amd import mlp
    flat = mlp.init_params(nin, nout, hidden, layers, seed=seed)
    rng = np.random.default_rng(seed + 1)
    for l, (W, b) in enumerate(mlp.unflatten(flat, nin, nout, hidden, layers)):
        b[: (nout if l == layers else hidden)] = rng.standard_normal(nout if l == layers else hidden) * 0.05
    net = mlp.MLP(nin, nout, hidden, layers)
    net.set_params(flat)
    x = np.zeros((B, net.in_pad), dtype=np.float32)
    x[:, :nin] = rng.standard_normal((B, nin))
    tgt = rng.standard_normal((B, nout)).astype(np.float32) * 0.3
    params = mlp.unflatten(flat.astype(np.float64), nin, nout, hidden, layers)
    return net, flat, params, x, tgt


def test_forward_matches_reference(torch_dev):
    torch = torch_dev
    net, _, params, x, _ = _net_case(torch)
    out = net.forward(_bf(torch, x))
    torch.cuda.synchronize()
    ref, _, _ = mlp_ref.forward(params, x, quantize=True)
    got = out.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 5e-3, err   # bf16 activations: a rounding flip changes one activation by 2^-8
    assert np.all(got[:, 37:] == 0.0)
    exact, _, _ = mlp_ref.forward(params, x, quantize=False)
    print("\nforward: max rel err vs bf16-emulating FP64 %.2e, vs exact FP64 %.2e"
          % (err, np.abs(got - exact).max() / np.abs(exact).max()))


@pytest.mark.parametrize("B", [128, 1024])
def test_gradients_match_reference(torch_dev, B):
    """Every parameter's gradient: relative L2 error per tensor against the
    FP64 reference with the same bf16 rounding points (tight), and against
    the exact FP64 gradient of the same parameters (loose: bf16 operands)."""
    torch = torch_dev
    from tiger_nga_amd import mlp
    net, flat, params, x, tgt = _net_case(torch, B=B)   # B = 1024: weight gradients split over the batch
    g, loss = net.gradients(_bf(torch, x), torch.from_numpy(tgt).cuda())
    lq, gq = mlp_ref.loss_and_grad(params, x, tgt, 37, quantize=True)
    le, ge = mlp_ref.loss_and_grad(params, x, tgt, 37, quantize=False)
    assert abs(loss - lq) <= 1e-3 * lq, (loss, lq)
    got = mlp.unflatten(g, 38, 37, 256, 4)
    worst_q = worst_e = 0.0
    for l, ((dW, db), (rW, rb), (eW, eb)) in enumerate(zip(got, gq, ge)):
        for a, r, e in ((dW, rW, eW), (db, rb, eb)):
            nq = np.linalg.norm(a - r) / np.linalg.norm(r)
            ne = np.linalg.norm(a - e) / np.linalg.norm(e)
            worst_q, worst_e = max(worst_q, nq), max(worst_e, ne)
            assert nq < 1e-2, (l, nq)
            assert ne < 5e-2, (l, ne)
    print("\ngradients: loss %.6e (ref %.6e); worst rel L2 err vs bf16-emulating ref %.2e, vs exact FP64 %.2e"
          % (loss, lq, worst_q, worst_e))


def test_training_reduces_loss(torch_dev):
    """Adam on a smooth 38 -> 37 map: the loss falls by more than 20x."""
    torch = torch_dev
    from tiger_nga_amd import mlp
    nin, nout = 38, 37
    net = mlp.MLP(nin, nout, 256, 4)
    net.set_params(mlp.init_params(nin, nout, 256, 4, seed=9))
    rng = np.random.default_rng(10)
    P = rng.standard_normal((nin, nout)) / np.sqrt(nin)
    B = 2048
    x = np.zeros((B, net.in_pad), dtype=np.float32)
    x[:, :nin] = rng.uniform(-1, 1, (B, nin))
    tgt = np.tanh(x[:, :nin] @ P * 2.0).astype(np.float32)
    xd, td = _bf(torch, x), torch.from_numpy(tgt).cuda()
    l0 = net.train_step(xd, td, 1e-3, want_loss=True)
    for _ in range(400):
        net.train_step(xd, td, 1e-3)
    l1 = net.train_step(xd, td, 1e-3, want_loss=True)
    print("\ntraining: loss %.3e -> %.3e after 400 Adam steps" % (l0, l1))
    assert l1 < l0 / 20
    # the trained parameters round-trip through the host
    p = net.get_params()
    assert np.all(np.isfinite(p)) and p.size == net.nparam
