"""Inverse planning on the GPU: the projected-gradient weight step (pbh_weight_step)
against its NumPy FP32 restatement, and the robust optimisation loop over a
robust-scenario batch (summed and worst-case objectives)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["sum", "worst"])
def test_weight_step_equals_fp32_restatement(mode):
    from pbdose import robust
    rng = np.random.default_rng(11)
    n, ns = 5000, 4
    w = rng.random(n).astype(np.float32)
    gs = [rng.standard_normal(n).astype(np.float32) for _ in range(ns)]
    J = np.array([3.0, 7.5, 7.5, 1.0])
    step, lo, hi = np.float32(0.37), np.float32(0.05), np.float32(0.9)
    dev = torch.device("cuda")
    wt = torch.as_tensor(w, device=dev)
    out = torch.empty_like(wt)
    robust.weight_step(wt, out, [torch.as_tensor(g, device=dev) for g in gs], torch.as_tensor(J, device=dev),
                       float(step), mode, float(lo), float(hi))
    if mode == "sum":
        g = gs[0].copy()
        for k in range(1, ns):
            g = (g + gs[k]).astype(np.float32)
    else:
        g = gs[int(np.argmax(J))]                     # first maximum: scenario 1
    ref = np.minimum(np.maximum((w - (step * g).astype(np.float32)).astype(np.float32), lo), hi)
    assert np.array_equal(out.cpu().numpy(), ref)
    # in place (w_new == w_old)
    robust.weight_step(wt, wt, [torch.as_tensor(g, device=dev) for g in gs], torch.as_tensor(J, device=dev),
                       float(step), mode, float(lo), float(hi))
    assert np.array_equal(wt.cpu().numpy(), ref)


@pytest.mark.parametrize("mode", ["sum", "worst"])
def test_robust_optimisation_decreases_objective(mode):
    """2 fields x 3 scenarios on a small grid, prescription in a central cube: 12
    iterations of projected gradient keep F non-increasing, lower it, and keep
    the weights in bounds; the final weights reproduce the final F."""
    from pbdose import model, robust
    print(f"case robust optimisation ({mode})", flush=True)
    g = model.DoseGrid(48, 40, 32, 0.25, 0.25, 0.25, -5.875, -4.875, 0.0)
    fa = robust.Field(g, model.SpotLattice(9, 7, 3, 0.5, 0.5, energies=np.array([90.0, 110.0, 130.0]),
                                           weights=np.random.default_rng(0).lognormal(0, 0.5, (3, 7, 9))))
    fb = robust.Field(g, model.SpotLattice(8, 8, 2, 0.5, 0.5, 0.25, 0.0, energies=np.array([100.0, 120.0]),
                                           weights=np.random.default_rng(1).lognormal(0, 0.5, (2, 8, 8))))
    S = robust.Scenario
    scen = [S(), S(0.3, 0, 0, 1.0), S(0, 0, 0, 0.965)]
    presc, vox = [], []
    for f in (fa, fb):
        Dp, wv = robust.central_cube(f.grid, half=1.5, dose_value=20.0)
        presc.append(Dp)
        vox.append(wv)
    b = robust.RobustBatch([fa, fb], scen, precision="bf16x3", presc=presc, vox_w=vox,
                           tuning=robust.DEFAULT_TUNING)
    h = robust.optimize(b, iters=12, mode=mode, w_min=0.0, w_max=50.0)
    F = np.array(h["F"])
    print(f"  F per field: start {F[0]}, end {F[-1]}; accepted {np.mean(h['accepted']):.2f}", flush=True)
    assert np.all(np.diff(F, axis=0) <= 0)
    assert np.all(F[-1] < 0.7 * F[0])
    for w in b.w:
        assert float(w.min()) >= 0.0 and float(w.max()) <= 50.0
    b.run()
    torch.cuda.synchronize()
    Jh = b.J.cpu().numpy().reshape(2, 3)
    Fr = Jh.sum(axis=1) if mode == "sum" else Jh.max(axis=1)
    assert np.allclose(Fr, F[-1], rtol=1e-12)
