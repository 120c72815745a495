"""finitechem_diffusion on the GPU (tgr_chem_diffusion_device, C ABI) against
the CPU oracle (oracle/diffusion.c, pinned in tests/test_diffusion.py), on
random fields over stretched grids with the reference's metric operators:
second and fourth order, molar diffusion on and off, gas radiation, cv, a
wall row in the mask, and the update of the transport source.  Outputs agree
to 1e-12 of each field's scale (the kernels sum the same terms in the same
order; FMA contraction changes the last bits)."""
import numpy as np
import pytest

from tiger_nga_amd import inputs, metric
from tiger_nga_amd.finitechem import Grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    return torch


def _dev(torch, a):
    strides = tuple(int(np.prod(a.shape[:k])) for k in range(a.ndim))
    dtype = torch.float64 if a.dtype == np.float64 else torch.int32
    t = torch.empty_strided(a.shape, strides, dtype=dtype, device="cuda:0")
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0"))
    return t


@pytest.mark.parametrize("mech,workload,order,molar,gasrad,cv,wall", [
    ("gri30", "gri30_uneven", 2, True, False, False, False), ("gri30", "gri30_uneven", 2, False, True, False, True),
    ("gri30", "gri30_uneven", 4, True, True, True, False), ("h2", "h2", 2, True, False, False, True),
    ("dme39", "dme39", 4, True, True, False, True)])
def test_diffusion_vs_oracle(torch_dev, oracle, mech, workload, order, molar, gasrad, cv, wall):
    torch = torch_dev
    from tiger_nga_amd.finitechem import FiniteChem
    nx, ny, nz = 9, 7, 6
    nover = {2: 1, 4: 3}[order]
    sh = (nx + 2 * nover, ny + 2 * nover, nz + 2 * nover)
    _, Y, T, rho, P, dt = inputs.states(workload, n=int(np.prod(sh)), seed=17)
    nsp = Y.shape[1]
    g = Grid.block(nx, ny, nz, nsp + 2, 1, nsp + 1, nover=nover)  # one extra scalar after T
    SC_ = np.zeros(sh + (nsp + 2,), order="F")
    for k in range(nsp):
        SC_[..., k] = Y[:, k].reshape(sh, order="F")
    SC_[..., nsp] = T.reshape(sh, order="F")
    SC_[..., nsp + 1] = 0.3
    RHO_ = np.asfortranarray(rho.reshape(sh, order="F"))
    rng = np.random.default_rng(5)
    DIFF_ = np.asfortranarray(10.0 ** rng.uniform(-5.0, -4.0, sh + (nsp + 2,)))
    x = np.cumsum(np.concatenate([[0.0], 1e-4 * (1.0 + 0.3 * rng.random(sh[0]))]))
    y = np.cumsum(np.concatenate([[0.0], 1e-4 * (1.0 + 0.3 * rng.random(sh[1]))]))
    m = metric.cartesian(g, x, y, 1.1e-4, order)
    mask = None
    if wall:
        mask = np.zeros(g.mask_shape, dtype=np.int32, order="F")
        mask[:, nover + 2] = 1  # a wall row at j = jmin_ + 2
    dc_o, gp_o = oracle.diffusion(mech, g, m, dt, RHO_, SC_, DIFF_, mask=mask, molar=molar, gasrad=gasrad, cv=cv)
    fc = FiniteChem(mech, 1e-6, 1e-12)
    md = {k: (v if k in ("st1", "st2") else _dev(torch, np.asfortranarray(v))) for k, v in m.items()}
    dc = _dev(torch, np.zeros(sh + (nsp + 1,), order="F"))
    gp = _dev(torch, np.zeros(sh + (3,), order="F"))
    src0 = np.asfortranarray(rng.standard_normal(SC_.shape))
    src = _dev(torch, src0)
    print("diffusion: %s order %d molar %d gasrad %d cv %d wall %d" % (mech, order, molar, gasrad, cv, wall),
          flush=True)
    fc.diffusion(dt, g, md, _dev(torch, RHO_), _dev(torch, SC_), _dev(torch, DIFF_),
                 mask=None if mask is None else _dev(torch, mask), diff_corr=dc, diff_gp=gp, src=src, molar=molar,
                 gasrad=gasrad, cv=cv)
    torch.cuda.synchronize()
    fc.close()
    dc, gp, src = dc.cpu().numpy(), gp.cpu().numpy(), src.cpu().numpy()
    for k in range(nsp + 1):
        scale = max(np.abs(dc_o[..., k]).max(), 1e-300)
        assert np.abs(dc[..., k] - dc_o[..., k]).max() <= 1e-12 * scale + 1e-300, k
    for q in range(3):
        scale = max(np.abs(gp_o[..., q]).max(), 1e-300)
        assert np.abs(gp[..., q] - gp_o[..., q]).max() <= 1e-12 * scale + 1e-300, q
    # finitechem_source_transport: src(species, T) += diff_corr on interior non-wall cells, nothing else
    want = src0.copy(order="F")
    want[..., :nsp + 1] += dc_o
    assert np.allclose(src, want, rtol=1e-12, atol=1e-12 * np.abs(dc_o).max())
    assert np.array_equal(src[..., nsp + 1], src0[..., nsp + 1])
    print("   max |diff_corr| species %.3e, T %.3e" % (np.abs(dc_o[..., :nsp]).max(), np.abs(dc_o[..., nsp]).max()))
