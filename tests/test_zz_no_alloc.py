"""The non-allocating device entry point (tgr_chem_source_chemistry_device_noalloc)
against the ordinary device entry point (tgr_chem_source_chemistry_device),
both called through the C ABI with ctypes.

The two run the same step kernel on the same cells; the non-allocating one
takes its work counter, counter totals and per-wave Jacobian workspace from
the library's file-scope device scratch.  Its results must equal the other
entry point's bit for bit, on two calls in a row (the second on a smaller
block: the scratch keeps the first call's counters, which the call must clear),
the call must leave the device's free memory unchanged (measured after the
ordinary entry point has run the same kernel once, so that the runtime's own
lazy allocations -- code object, register-spill backing -- are already made),
and a grid whose workspace exceeds the scratch must be refused on the host
before any launch.
"""
import ctypes

import numpy as np
import pytest

from tiger_nga_amd import inputs
from tiger_nga_amd.finitechem import FiniteChem, Grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    return torch


def _fortran_dev(torch, a):
    """Device copy of a numpy array with Fortran strides (element order kept)."""
    strides = tuple(int(np.prod(a.shape[:k])) for k in range(a.ndim))
    dtype = torch.float64 if a.dtype == np.float64 else torch.int32
    t = torch.empty_strided(a.shape, strides, dtype=dtype, device="cuda:0")
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0"))
    return t


def _case(torch, workload, nx, seed):
    mech, Y, T, rho, P, dt = inputs.states(workload, n=nx ** 3, seed=seed)
    RHO_, SC_, isc_sc, isc_T = inputs.fortran_fields(Y, T, rho, nx, nx, nx, nover=1)
    g = Grid.block(nx, nx, nx, SC_.shape[3], isc_sc, isc_T, nover=1)
    mask = np.zeros(g.mask_shape, dtype=np.int32, order="F")
    mask[1, :] = 1  # one wall row
    return mech, P, dt, g, RHO_, SC_, mask


def _run(torch, fc, entry, g, RHO_, SC_, mask, P, dt):
    """One call of `entry` on fresh device copies; returns (SC_, HR_gp, chemsrc_gp) on the host."""
    R = _fortran_dev(torch, RHO_)
    S = _fortran_dev(torch, SC_)
    M = _fortran_dev(torch, mask)
    HR = _fortran_dev(torch, np.full(RHO_.shape, 7.0, order="F"))
    CS = _fortran_dev(torch, np.full(RHO_.shape + (3,), 7.0, order="F"))
    torch.cuda.synchronize()
    gc = g._c()
    args = (fc._h, dt, P, P, ctypes.byref(gc), R.data_ptr(), S.data_ptr(), M.data_ptr(), HR.data_ptr(),
            CS.data_ptr(), None)
    free0 = torch.cuda.mem_get_info()[0]
    if entry == "noalloc":
        rc = fc.lib.tgr_chem_source_chemistry_device_noalloc(*args)
    else:
        rc = fc.lib.tgr_chem_source_chemistry_device(*args, None)
    assert rc == 0, fc.lib.tgr_chem_last_error().decode()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    return S.cpu().numpy(), HR.cpu().numpy(), CS.cpu().numpy(), free0 - free1


def test_noalloc_matches_device_entry(torch_dev):
    torch = torch_dev
    fc = FiniteChem("gri30", 1e-6, 1e-12)
    cases = []
    for nx, seed in ((12, 31), (5, 32)):
        mech, P, dt, g, RHO_, SC_, mask = _case(torch, "gri30_uneven", nx, seed)
        print("device entry: gri30_uneven block %d^3, seed %d" % (nx, seed), flush=True)
        cases.append(((g, RHO_, SC_, mask, P, dt), _run(torch, fc, "device", g, RHO_, SC_, mask, P, dt)))
    # two non-allocating calls in a row, the second on the smaller block
    for (args, ref), (nx, seed) in zip(cases, ((12, 31), (5, 32))):
        print("non-allocating entry: gri30_uneven block %d^3, seed %d" % (nx, seed), flush=True)
        Sn, Hn, Cn, used = _run(torch, fc, "noalloc", *args)
        assert used == 0, "the non-allocating entry point changed the free device memory by %d bytes" % used
        assert not np.array_equal(Sn, args[2])  # the step did something
        assert np.array_equal(Sn, ref[0]) and np.array_equal(Hn, ref[1]) and np.array_equal(Cn, ref[2])
    fc.close()


def test_noalloc_refuses_oversized_workspace(torch_dev):
    torch = torch_dev
    fc = FiniteChem("gri30", 1e-6, 1e-12)
    nx = 32
    mech, P, dt, g, RHO_, SC_, mask = _case(torch, "gri30", nx, 33)
    fc.set_option("grid", 4096)  # 4096 workgroups x 8 waves x 37 x 512 B = 620 MB of workspace
    R = _fortran_dev(torch, RHO_)
    S = _fortran_dev(torch, SC_)
    before = S.cpu().numpy()
    gc = g._c()
    rc = fc.lib.tgr_chem_source_chemistry_device_noalloc(fc._h, dt, P, P, ctypes.byref(gc), R.data_ptr(),
                                                         S.data_ptr(), None, None, None, None)
    torch.cuda.synchronize()
    assert rc == -1 and b"device scratch" in fc.lib.tgr_chem_last_error()
    assert np.array_equal(S.cpu().numpy(), before)
    fc.close()
