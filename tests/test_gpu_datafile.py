"""The chemistry step on fields read from an NGA data file: the file is
loaded straight into device memory, stepped by the HIP path through the C
ABI, written back, and compared with the oracle stepping the same file's
fields on the host."""
import numpy as np
import pytest

from tiger_nga_amd import ngadata
from tiger_nga_amd.mech import meta

from .test_ngadata import _ignition_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mech", ["h2", "gri30"])
def test_step_on_data_file(mech, oracle, tmp_path):
    import torch
    from tiger_nga_amd.finitechem import FiniteChem
    assert torch.cuda.is_available(), "GPU test without a HIP device"
    rtol, atol = 1e-10, 1e-16
    p = str(tmp_path / "in.dat")
    _, P, dt = _ignition_file(p, mech, 6, 5, 4, 31)
    df = ngadata.read(p, expect_shape=(6, 5, 4))
    species = meta.species(mech)
    R_d, S_d, g = ngadata.chemistry_block(df, species, nover=2, device="cuda")
    fc = FiniteChem(mech, rtol, atol)
    st = fc.source_chemistry(R_d, S_d, g, dt, P, P)
    torch.cuda.synchronize()
    RHO_, SC_o, go = ngadata.chemistry_block(df, species, nover=2)
    assert (go.isc_sc, go.isc_T, go.nscalar) == (g.isc_sc, g.isc_T, g.nscalar)
    sto = oracle.source_chemistry(mech, RHO_, SC_o, go, dt, P, P, rtol=rtol, atol=atol)
    assert st["failed"] == 0 and sto["status"] == 0 and st["cells"] == 6 * 5 * 4
    q = str(tmp_path / "out.dat")
    ngadata.write_chemistry_block(q, df, R_d, S_d, g, time=df.time + dt)
    out = ngadata.read(q, expect_shape=(6, 5, 4))
    assert out.names == df.names and out.time == df.time + dt
    nsp = len(species)
    sc = out.sc_names
    for k, s in enumerate(species):
        got = np.asarray(out.var(s))
        ref = SC_o[2:8, 2:7, 2:6, g.isc_sc - 1 + k]
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-12)
        assert err.max() < 1e-8, "%s: max rel err %.3e" % (s, err.max())
    dT = np.abs(np.asarray(out.var("T")) - SC_o[2:8, 2:7, 2:6, g.isc_T - 1]).max()
    assert dT < 1e-6
    # untouched: flow variables, ZMIX, and the density
    for n in ("U", "V", "W", "P", "RHO", "dRHO", "ZMIX"):
        assert np.array_equal(np.asarray(out.var(n)), np.asarray(df.var(n))), n
    assert sc[g.isc_sc - 1:g.isc_sc - 1 + nsp] == list(species)
    print("\n%s from a data file: %d cells, steps %d, max |dT| %.2e K" % (mech, st["cells"], st["nst"], dT))
