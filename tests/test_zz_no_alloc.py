"""pbh_forward_noalloc (forward dose with its workspace in the library's fixed
device scratch) through the C ABI, against pbh_forward_tuned with a caller
workspace: bit-equal, on two calls in a row, the second on a smaller problem
(the scratch keeps the first call's contents). Also the refusal of a problem
that needs more scratch than the fixed size, before any launch."""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz, d, nsx, nsy, energies, seed):
    from pbdose import model
    g = model.DoseGrid(nx, ny, nz, d, d, d, -(nx - 1) * d / 2, -(ny - 1) * d / 2, 0.0)
    E = np.asarray(energies, dtype=np.float64)
    w = np.random.default_rng(seed).lognormal(0.0, 0.5, (len(E), nsy, nsx))
    idd, sig = O.layer_luts(g.z, E)
    lat = model.SpotLattice(nsx, nsy, len(E), 0.5, 0.5, energies=E, weights=w)
    return model.PBProblem(g, lat, idd, sig)


def _noalloc(dp, precision, chunk):
    from pbdose import _lib
    out = torch.full(dp.shape, float("nan"), dtype=torch.float32, device=dp.dev)
    rc = _lib.lib().pbh_forward_noalloc(dp.grid_s, dp.lat_s, _lib.ptr(dp.w), _lib.ptr(dp.idd),
                                        _lib.ptr(dp.sigma), _lib.ptr(out), _lib.PREC[precision],
                                        int(chunk), _lib.stream_of(dp.dev))
    _lib.check(rc, "pbh_forward_noalloc")
    return out


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_noalloc_forward_equals_workspace_forward(precision):
    from pbdose import _lib, engines
    big = _problem(128, 128, 128, 0.2, 41, 41, np.linspace(70.0, 180.0, 40), 0)
    small = _problem(40, 36, 24, 0.25, 11, 9, (80.0, 110.0, 140.0), 3)
    for name, p in (("128^3 / 41x41x40", big), ("40x36x24 / 11x9x3", small)):
        print(f"case noalloc {precision} {name}", flush=True)
        dp = engines.DeviceProblem(p)
        chunk = dp.auto_chunk(_lib.PREC[precision])
        need = _lib.lib().pbh_forward_workspace_bytes_tuned(dp.grid_s, dp.lat_s, _lib.PREC[precision], chunk)
        assert 0 < need <= _lib.PBH_NOALLOC_SCRATCH_BYTES
        ref = dp.forward(precision, layer_chunk=chunk)
        got = _noalloc(dp, precision, chunk)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"max diff {float((got - ref).abs().max()):.3e}"
        assert float(ref.abs().max()) > 0


def test_noalloc_refuses_oversized_problems_before_launching():
    from pbdose import _lib, engines
    p = _problem(128, 128, 128, 0.2, 41, 41, np.linspace(70.0, 180.0, 40), 0)
    dp = engines.DeviceProblem(p)
    # one layer per chunk: 39 partial dose volumes, far beyond the fixed scratch
    need = _lib.lib().pbh_forward_workspace_bytes_tuned(dp.grid_s, dp.lat_s, 3, 1)
    assert need > _lib.PBH_NOALLOC_SCRATCH_BYTES
    out = torch.zeros(dp.shape, dtype=torch.float32, device=dp.dev)
    rc = _lib.lib().pbh_forward_noalloc(dp.grid_s, dp.lat_s, _lib.ptr(dp.w), _lib.ptr(dp.idd),
                                        _lib.ptr(dp.sigma), _lib.ptr(out), 3, 1, _lib.stream_of(dp.dev))
    assert rc == -8 and "scratch" in _lib.strerror(rc)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
