"""The committed trained networks (tiger_nga_amd/models/*.bin) read back
with a consistent shape for their mechanism, zero padding and finite values,
and a damaged file is refused."""
import numpy as np
import pytest

from tiger_nga_amd import mlp, surrogate
from tiger_nga_amd.mech import meta


@pytest.mark.parametrize("mech", ["h2", "gri30", "dme39", "ethyred"])
def test_model_file(mech):
    (nin, nout, hidden, layers), shift, scale, oscale, params, info = surrogate.Surrogate.read(
        surrogate.model_path(mech))
    nsp = len(meta.species(mech))
    assert (nin, nout) == (nsp + 2, nsp + 1)
    assert np.all(np.isfinite(shift)) and np.all(np.isfinite(scale)) and np.all(scale > 0)
    assert np.all(np.isfinite(oscale)) and np.all(oscale > 0)
    assert np.all(np.isfinite(params))
    for l, (W, b) in enumerate(mlp.unflatten(params, nin, nout, hidden, layers)):
        fo, fi, no, ni = mlp.layer_shapes(nin, nout, hidden, layers)[l]
        assert np.all(W[no:] == 0) and np.all(W[:, ni:] == 0) and np.all(b[no:] == 0)
    assert info.get("mech") == mech and info.get("pairs", 0) > 1_000_000


def test_damaged_model_file_is_refused(tmp_path):
    raw = open(surrogate.model_path("h2"), "rb").read()
    p = tmp_path / "cut.bin"
    p.write_bytes(raw[:-4])
    with pytest.raises(ValueError, match="size"):
        surrogate.Surrogate.read(str(p))
    q = tmp_path / "magic.bin"
    q.write_bytes(b"XXXXXXXX" + raw[8:])
    with pytest.raises(ValueError, match="not a network"):
        surrogate.Surrogate.read(str(q))
