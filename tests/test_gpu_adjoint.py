"""Hand-written adjoint (GPU, through the C ABI) vs the FP64 oracle, central
finite differences and autograd through a PyTorch restatement of the reference
engine.

Tolerances (FP32 kernels against FP64 references, relative to the largest
magnitude of the compared gradient): 2e-5 for dJ/dw and dJ/dIDD, 5e-5 for
dJ/dsigma and the range-shift gradient.
"""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu


def small_problem(seed=3, nx=40, ny=36, nz=24, d=0.25, nsx=11, nsy=9, ds=0.5, energies=(80.0, 110.0, 140.0)):
    from pbdose import model
    g = model.DoseGrid(nx, ny, nz, d, d, d, -(nx - 1) * d / 2, -(ny - 1) * d / 2, 0.0)
    E = np.asarray(energies, dtype=np.float64)
    w = np.random.default_rng(seed).lognormal(0.0, 0.5, (len(E), nsy, nsx))
    idd, sig = O.layer_luts(g.z, E)
    lat = model.SpotLattice(nsx, nsy, len(E), ds, ds, energies=E, weights=w)
    return model.PBProblem(g, lat, idd, sig)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_dot_product_identity():
    from pbdose import engines, planning
    print("case adjoint dot-product test", flush=True)
    p = small_problem()
    dp = engines.DeviceProblem(p)
    torch.manual_seed(0)
    R = torch.randn(dp.shape, device=dp.dev, dtype=torch.float32)
    D = dp.forward()
    gw = planning.adjoint(dp, R)
    lhs = float((R.double() * D.double()).sum())
    rhs = float((gw.double() * dp.w.double()).sum())
    print(f"  <R, D w> = {lhs:.9e}  <D^T R, w> = {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("shape", [(40, 36, 24, 11, 9), (70, 150, 6, 60, 50), (130, 65, 5, 7, 13)])
def test_adjoint_matches_oracle(shape):
    from pbdose import engines, planning
    nx, ny, nz, nsx, nsy = shape
    print(f"case adjoint vs oracle {shape}", flush=True)
    p = small_problem(nx=nx, ny=ny, nz=nz, nsx=nsx, nsy=nsy, ds=0.3)
    dp = engines.DeviceProblem(p)
    R = np.random.default_rng(1).standard_normal(p.grid.shape)
    gw, gi, gs = planning.adjoint(dp, torch.as_tensor(R, dtype=torch.float32, device=dp.dev), lut_grads=True)
    x, y, sx, sy = p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y
    Rf = R.astype(np.float32).astype(np.float64)
    ref_w = O.adjoint_w(x, y, sx, sy, p.idd, p.sigma, Rf)
    ref_i, ref_s = O.adjoint_lut(x, y, sx, sy, p.lattice.weights, p.idd, p.sigma, Rf)
    ew, ei, es = rel(gw.cpu(), ref_w), rel(gi.cpu(), ref_i), rel(gs.cpu(), ref_s)
    print(f"  dJ/dw {ew:.2e}  dJ/dIDD {ei:.2e}  dJ/dsigma {es:.2e}", flush=True)
    assert ew <= 2e-5 and ei <= 2e-5 and es <= 5e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_adjoint_two_to_three_pitch(precision):
    """2 mm voxels and 3 mm spots: parity-interleaved tables with P = 2, Q = 3 (run-time
    slots, odd entries at the bank offset). FP32 with table gradients; BF16x3 dJ/dw
    within the mixed adjoint's bound."""
    from pbdose import engines, planning
    print(f"case adjoint 2 mm / 3 mm {precision}", flush=True)
    p = small_problem(nx=80, ny=56, nz=20, d=0.2, nsx=23, nsy=15, ds=0.3)
    dp = engines.DeviceProblem(p)
    R = np.random.default_rng(6).standard_normal(p.grid.shape)
    Rt = torch.as_tensor(R, dtype=torch.float32, device=dp.dev)
    x, y, sx, sy = p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y
    Rf = R.astype(np.float32).astype(np.float64)
    ref_w = O.adjoint_w(x, y, sx, sy, p.idd, p.sigma, Rf)
    if precision == "fp32":
        gw, gi, gs = planning.adjoint(dp, Rt, lut_grads=True)
        ref_i, ref_s = O.adjoint_lut(x, y, sx, sy, p.lattice.weights, p.idd, p.sigma, Rf)
        ew, ei, es = rel(gw.cpu(), ref_w), rel(gi.cpu(), ref_i), rel(gs.cpu(), ref_s)
        print(f"  dJ/dw {ew:.2e}  dJ/dIDD {ei:.2e}  dJ/dsigma {es:.2e}", flush=True)
        assert ew <= 2e-5 and ei <= 2e-5 and es <= 5e-5
    else:
        gw = planning.adjoint(dp, Rt, precision=precision)
        ew = rel(gw.cpu(), ref_w)
        print(f"  dJ/dw {ew:.2e}", flush=True)
        assert ew <= MIXED_ADJ_TOL[precision]


def test_adjoint_direct_factor_path():
    from pbdose import engines, planning
    print("case adjoint with direct factors (0.37 cm spots)", flush=True)
    p = small_problem(nsx=13, nsy=7, ds=0.37)
    dp = engines.DeviceProblem(p)
    R = np.random.default_rng(2).standard_normal(p.grid.shape)
    gw = planning.adjoint(dp, torch.as_tensor(R, dtype=torch.float32, device=dp.dev))
    ref = O.adjoint_w(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y, p.idd, p.sigma,
                      R.astype(np.float32).astype(np.float64))
    assert rel(gw.cpu(), ref) <= 2e-5


def test_objective_kernel():
    from pbdose import engines, planning
    p = small_problem()
    dp = engines.DeviceProblem(p)
    D = dp.forward()
    Dp, wv = planning.central_cube(p.grid, half=1.5, dose_value=float(D.max()) * 0.5)
    J, R = planning.objective(dp, D, Dp, wv)
    d64 = D.double() - Dp.double()
    assert abs(float(J) - float((wv.double() * d64 * d64).sum())) <= 1e-9 * float(J)
    assert torch.allclose(R.double(), 2 * wv.double() * d64, rtol=1e-6, atol=1e-6 * float(R.abs().max()))
    # the objective's workspace is kept with the problem: repeated calls give the same
    # bits, and a call with other inputs its own J
    for _ in range(3):
        J2, _ = planning.objective(dp, D, Dp, wv)
        assert float(J2) == float(J)
    Dp3 = Dp * 0.5
    J3, _ = planning.objective(dp, D, Dp3, None)
    d3 = D.double() - Dp3.double()
    assert abs(float(J3) - float((d3 * d3).sum())) <= 1e-9 * float(J3)


def test_clinical_objective_gradients_vs_finite_differences():
    """128^3 / 41x41 / 40-layer field, J = sum wv (D - Dp)^2 with Dp = 2 in the
    central 6 cm cube, wv = 1 inside / 0.1 outside; central differences with
    h = 1e-3 on 64 random spots (FP64 oracle dose) and on 6 layers' range shifts."""
    from pbdose import model, planning
    print("case C4: objective gradients on the 128^3 field", flush=True)
    grid = model.DoseGrid.centered(128, 0.2)
    E = np.linspace(70.0, 180.0, 40)
    w = np.random.default_rng(seed=0).lognormal(0.0, 0.5, (40, 41, 41))
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5, energies=E, weights=w)
    fm = planning.FieldModel(grid, lat)
    Dp_t, wv_t = planning.central_cube(grid)
    wt = torch.as_tensor(w, dtype=torch.float32, device=fm.dev).requires_grad_(True)
    delta = torch.zeros(40, dtype=torch.float64, device=fm.dev, requires_grad=True)
    D = planning.dose(wt, fm, delta)
    J = (wv_t.double() * (D.double() - Dp_t.double()) ** 2).sum()
    J.backward()
    gw = wt.grad.double().cpu().numpy()
    gd = delta.grad.cpu().numpy()

    # FP64 oracle: dose, objective, and exact central differences
    x, y, z = grid.x, grid.y, grid.z
    sx = lat.spot_x
    idd, sig = O.layer_luts(z, E)
    w32 = w.astype(np.float32).astype(np.float64)
    Dref = O.separable_dose(x, y, sx, sx, w32, idd, sig)
    Dp, wv = O.presc_cube(O.centered_grid(128, 0.2))
    rng = np.random.default_rng(1)
    spots = [(int(rng.integers(40)), int(rng.integers(41)), int(rng.integers(41))) for _ in range(64)]
    h = 1e-3
    fd_w, g_w = [], []
    for (l, j, i) in spots:
        S = O.spot_dose(x, y, sx[i], sx[j], idd[l], sig[l])
        fd = (O.objective(Dref + h * S, Dp, wv) - O.objective(Dref - h * S, Dp, wv)) / (2 * h)
        fd_w.append(fd)
        g_w.append(gw[l, j, i])
    ew = rel(g_w, fd_w)
    # range shifts: the FP64 chain rule through the device's tangent tables (checks
    # the adjoint's table gradients and the depth contraction), and central
    # differences of the FP64 objective
    R = 2.0 * wv * (Dref - Dp)
    didd_t, dsig_t = (t.cpu().numpy() for t in fm.lut_derivatives(None))
    ch_d, g_d, fd_d = [], [], {h: [] for h in (1e-3, 2e-6, 1e-6)}
    for l in (0, 9, 17, 25, 33, 39):
        Dl_old = O.separable_dose(x, y, sx, sx, w32[l:l + 1], idd[l:l + 1], sig[l:l + 1])
        for hh in fd_d:
            parts = []
            for sgn in (+1, -1):
                i_s, s_s = O.layer_luts(z, O.shifted_energies(E[l:l + 1], [sgn * hh]))
                Dl_new = O.separable_dose(x, y, sx, sx, w32[l:l + 1], i_s, s_s)
                parts.append(O.objective(Dref - Dl_old + Dl_new, Dp, wv))
            fd_d[hh].append((parts[0] - parts[1]) / (2 * hh))
        gi, gs = O.adjoint_lut(x, y, sx, sx, w32[l:l + 1], idd[l:l + 1], sig[l:l + 1], R)
        ch_d.append(float(np.sum(gi * didd_t[l:l + 1] + gs * dsig_t[l:l + 1])))
        g_d.append(gd[l])
    rich = [(4.0 * a - b) / 3.0 for a, b in zip(fd_d[1e-6], fd_d[2e-6])]
    ec, er, e3 = rel(g_d, ch_d), rel(g_d, rich), rel(g_d, fd_d[1e-3])
    print(f"  dJ/dw vs FD (64 spots) {ew:.2e}; dJ/ddelta vs FP64 chain rule {ec:.2e}, "
          f"vs FD h->0 (Richardson of 2e-6, 1e-6) {er:.2e}, vs FD h = 1e-3 {e3:.2e}", flush=True)
    assert ew <= 2e-5
    assert ec <= 5e-5
    # J(delta) is piecewise smooth with dense breaks (range-table knots crossing
    # CSDA bin edges, fine-grid points crossing R0, physics.py:255-264, :359-370),
    # some 1e-5 cm apart, so differences over +-1e-3 cm average the slope over
    # many pieces: the analytic (local) slope departs from that average by a few
    # per cent (measured 2.6e-2) while it equals the limit h -> 0 (measured 2.9e-7)
    assert er <= 1e-5
    assert e3 <= 5e-2


def test_autograd_matches_torch_restatement_of_reference_engine():
    """Gradients from the hand-written adjoint vs torch autograd through the
    batched-matmul engine (oracle/torch_cpu.py, FP64, weights as leaf)."""
    from oracle.torch_cpu import torch_separable_cpu
    from pbdose import engines, planning
    p = small_problem()
    dp = engines.DeviceProblem(p)
    Dp_np = np.zeros(p.grid.shape)
    Dp_np[8:16, 12:24, 14:26] = 30.0
    wv_np = np.where(Dp_np > 0, 1.0, 0.1)
    wl = torch.tensor(p.lattice.weights.astype(np.float32).astype(np.float64), requires_grad=True)
    Dt = torch_separable_cpu(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y, wl, p.idd, p.sigma)
    Jt = (torch.as_tensor(wv_np) * (Dt - torch.as_tensor(Dp_np)) ** 2).sum()
    Jt.backward()
    D = dp.forward()
    _, R = planning.objective(dp, D, torch.as_tensor(Dp_np, dtype=torch.float32, device=dp.dev),
                              torch.as_tensor(wv_np, dtype=torch.float32, device=dp.dev))
    gw = planning.adjoint(dp, R)
    e = rel(gw.cpu(), wl.grad.numpy())
    print(f"  hand-written adjoint vs torch autograd (FP64): {e:.2e}", flush=True)
    assert e <= 2e-5


@pytest.mark.parametrize("adj", ["fp16x3", "bf16x3"])
def test_dose_autograd_with_16bit_adjoint(adj):
    """planning.dose(..., adjoint_precision) backward: dJ/dw from the 16-bit adjoint
    vs torch autograd through the batched-matmul engine (FP64); with a range-shift
    leaf the backward runs the FP32 adjoint (table gradients are FP32 only), so
    dJ/dw then equals the FP32 backward's bit for bit."""
    from oracle.torch_cpu import torch_separable_cpu
    from pbdose import planning
    print(f"case dose autograd with the {adj} adjoint", flush=True)
    fm = planning.FieldModel(small_problem().grid, small_problem().lattice)
    p = fm.problem   # the field's own device-built tables
    Dp_np = np.zeros(p.grid.shape)
    Dp_np[8:16, 12:24, 14:26] = 30.0
    wv_np = np.where(Dp_np > 0, 1.0, 0.1)
    wl = torch.tensor(p.lattice.weights.astype(np.float32).astype(np.float64), requires_grad=True)
    Dt = torch_separable_cpu(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y, wl, p.idd, p.sigma)
    (torch.as_tensor(wv_np) * (Dt - torch.as_tensor(Dp_np)) ** 2).sum().backward()
    Dp_t = torch.as_tensor(Dp_np, dtype=torch.float32, device=fm.dev)
    wv_t = torch.as_tensor(wv_np, dtype=torch.float32, device=fm.dev)

    def grads(adjoint_precision, with_delta):
        wt = torch.as_tensor(p.lattice.weights, dtype=torch.float32, device=fm.dev).requires_grad_(True)
        delta = (torch.zeros(p.lattice.n_layers, dtype=torch.float64, device=fm.dev, requires_grad=True)
                 if with_delta else None)
        D = planning.dose(wt, fm, delta, adjoint_precision=adjoint_precision)
        (wv_t * (D - Dp_t) ** 2).sum().backward()
        return wt.grad.clone(), (delta.grad.clone() if with_delta else None)

    g16, _ = grads(adj, False)
    e = rel(g16.cpu(), wl.grad.numpy())
    print(f"  dose backward with the {adj} adjoint vs torch autograd (FP64): {e:.2e}", flush=True)
    assert e <= {"fp16x3": 5e-6, "bf16x3": 3e-5}[adj]
    g32, d32 = grads("fp32", True)
    g16d, d16 = grads(adj, True)
    assert torch.equal(g16d, g32) and torch.equal(d16, d32)
    with pytest.raises(ValueError):
        planning.dose(torch.zeros(1, dtype=torch.float32, device=fm.dev), fm, adjoint_precision="fp64")


MIXED_ADJ_TOL = {"bf16x3": 3e-5, "bf16": 1e-2, "fp16x3": 5e-6, "fp16": 3e-3}


@pytest.mark.parametrize("precision", ["bf16x3", "bf16", "fp16x3", "fp16"])
@pytest.mark.parametrize("geom", [(40, 36, 24, 11, 9, 0.2, 0.5), (70, 150, 6, 60, 50, 0.25, 0.3),
                                  (130, 65, 5, 7, 13, 0.2, 0.5)])
def test_mixed_adjoint_matches_oracle(precision, geom):
    """16-bit MFMA adjoint (dJ/dw): 2 mm / 5 mm pitch (compile-time table slots),
    2.5 mm / 3 mm (run-time slots), ragged tiles and multi-chunk lattices."""
    from pbdose import engines, planning
    nx, ny, nz, nsx, nsy, d, ds = geom
    print(f"case mixed adjoint {precision} {geom}", flush=True)
    p = small_problem(nx=nx, ny=ny, nz=nz, d=d, nsx=nsx, nsy=nsy, ds=ds)
    dp = engines.DeviceProblem(p)
    R = np.random.default_rng(4).standard_normal(p.grid.shape)
    Rt = torch.as_tensor(R, dtype=torch.float32, device=dp.dev)
    gw = planning.adjoint(dp, Rt, precision=precision)
    ref = O.adjoint_w(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y, p.idd, p.sigma,
                      R.astype(np.float32).astype(np.float64))
    e = rel(gw.cpu(), ref)
    g2 = planning.adjoint(dp, Rt, precision=precision).clone()
    assert torch.equal(gw, g2), "mixed adjoint not deterministic"
    print(f"  dJ/dw vs oracle {e:.2e}", flush=True)
    assert e <= MIXED_ADJ_TOL[precision]


@pytest.mark.parametrize("precision", ["fp16x3", "fp16", "bf16x3"])
def test_16bit_adjoint_at_any_residual_scale(precision):
    """R scaled by 2^k gives dJ/dw scaled by 2^k, bit for bit: FP16 scales each wave's
    rows of R by a power of two and undoes it exactly. Against the FP32 adjoint: R
    whose rows 16-31 of every 64-row tile (one wave's rows) are 32 times larger (waves
    with different scales, brought to one common scale in stage 1), and R whose first
    or last tile group is 2^30 times larger (a scale per group)."""
    from pbdose import engines, planning
    print(f"case 16-bit adjoint at any residual scale {precision}", flush=True)
    p = small_problem(nx=64, ny=150, nz=12, d=0.2, nsx=21, nsy=17, ds=0.5,
                      energies=tuple(np.linspace(70.0, 150.0, 8)))
    dp = engines.DeviceProblem(p)
    R = torch.as_tensor(np.random.default_rng(9).standard_normal(p.grid.shape), dtype=torch.float32,
                        device=dp.dev)
    g0 = planning.adjoint(dp, R, precision=precision)
    for k in (-40, -12, 17, 40):
        gk = planning.adjoint(dp, R * 2.0 ** k, precision=precision)
        assert torch.equal(gk, g0 * 2.0 ** k), k
    wave1 = [y for y in range(p.grid.ny) if 16 <= y % 64 < 32]
    for name, rows, f in (("one wave's rows x 32", wave1, 32.0),
                          ("tiles 0-1 x 2^30", list(range(0, 128)), 2.0 ** 30),
                          ("tile 2 x 2^30", list(range(128, p.grid.ny)), 2.0 ** 30)):
        Rm = R.clone()
        Rm[:, rows, :] *= f
        g32 = planning.adjoint(dp, Rm)
        e = rel(planning.adjoint(dp, Rm, precision=precision).cpu(), g32.cpu())
        print(f"  {name}: vs FP32 adjoint {e:.2e}", flush=True)
        assert e <= MIXED_ADJ_TOL[precision], name


def test_mixed_adjoint_clinical_field_and_unsupported_cases():
    """128^3 / 41x41 / 40 layers, R of the weighted-quadratic objective: the BF16x3
    adjoint against the FP32 adjoint (itself within 2e-7 of the oracle); table
    gradients and direct-factor geometries refuse the BF16 form."""
    from pbdose import engines, model, physics, planning
    print("case mixed adjoint on the 128^3 field", flush=True)
    grid = model.DoseGrid.centered(128, 0.2)
    E = np.linspace(70.0, 180.0, 40)
    w = np.random.default_rng(seed=0).lognormal(0.0, 0.5, (40, 41, 41))
    idd, sig, _ = physics.build_luts(grid.nz, grid.oz, grid.dz, E)
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5, energies=E, weights=w)
    dp = engines.DeviceProblem(model.PBProblem(grid, lat, idd.cpu().numpy(), sig.cpu().numpy()), idd=idd, sigma=sig)
    Dp, wv = planning.central_cube(grid)
    _, R = planning.objective(dp, dp.forward("bf16x3"), Dp, wv)
    g32 = planning.adjoint(dp, R)
    for prec in ("bf16x3", "bf16", "fp16x3", "fp16"):
        e = rel(planning.adjoint(dp, R, precision=prec).cpu(), g32.cpu())
        print(f"  {prec} vs FP32 adjoint: {e:.2e}", flush=True)
        assert e <= MIXED_ADJ_TOL[prec]
    with pytest.raises(ValueError):
        planning.adjoint(dp, R, lut_grads=True, precision="bf16x3")
    p = small_problem(nsx=13, nsy=7, ds=0.37)
    dq = engines.DeviceProblem(p)
    with pytest.raises(RuntimeError, match="unsupported"):
        planning.adjoint(dq, torch.zeros(dq.shape, dtype=torch.float32, device=dq.dev), precision="bf16x3")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16", "fp16x3", "fp16"])
def test_adjoint_does_not_depend_on_the_split(precision):
    """dJ/dw is the same, bit for bit, for every layer split 1-16 (powers of two or
    not: each (slice, layer) partial is computed by one workgroup in a fixed order,
    and the depth reduction adds them in increasing z), and the automatic split is
    the chooser's."""
    from pbdose import _lib, engines, planning
    print(f"case adjoint split independence {precision}", flush=True)
    p = small_problem(nx=64, ny=48, nz=60, d=0.2, nsx=13, nsy=11, ds=0.5,
                      energies=tuple(np.linspace(70.0, 150.0, 12)))
    dp = engines.DeviceProblem(p)
    R = torch.as_tensor(np.random.default_rng(4).standard_normal(p.grid.shape), dtype=torch.float32,
                        device=dp.dev)
    ref = planning.adjoint(dp, R, split=1, precision=precision)
    for split in (2, 3, 5, 7, 8, 10, 16):
        g = planning.adjoint(dp, R, split=split, precision=precision)
        assert torch.equal(g, ref), f"split {split}"
    auto = dp.auto_split(_lib.PREC[precision])
    assert 1 <= auto <= 16
    assert torch.equal(planning.adjoint(dp, R, precision=precision), ref)
    if precision == "fp32":
        gw1, gi1, gs1 = planning.adjoint(dp, R, split=1, lut_grads=True)
        gw5, gi5, gs5 = planning.adjoint(dp, R, split=5, lut_grads=True)
        assert torch.equal(gw5, gw1) and torch.equal(gi5, gi1) and torch.equal(gs5, gs1)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3"])
def test_objective_gradient_equals_objective_then_adjoint(precision):
    """pbh_objective_adjoint gives J and dJ/dw bit for bit as pbh_objective followed
    by pbh_adjoint_tuned, on repeated calls (the workspace is reused), with and
    without voxel weights."""
    from pbdose import engines, planning
    print(f"case fused objective + adjoint {precision}", flush=True)
    p = small_problem(nx=64, ny=48, nz=40, d=0.2, nsx=13, nsy=11, ds=0.5,
                      energies=tuple(np.linspace(70.0, 130.0, 8)))
    dp = engines.DeviceProblem(p)
    D = dp.forward()
    Dp, wv = planning.central_cube(p.grid, half=1.5, dose_value=float(D.max()) * 0.5)
    for vw in (wv, None):
        J0, R = planning.objective(dp, D, Dp, vw)
        g0 = planning.adjoint(dp, R, precision=precision)
        for _ in range(2):
            J1, g1 = planning.objective_gradient(dp, D, Dp, vw, precision=precision)
            assert float(J1) == float(J0)
            assert torch.equal(g1, g0)


@pytest.mark.parametrize("fwd,adj,rows", [("fp32", "fp32", 32), ("bf16x3", "bf16x3", 32), ("bf16x3", "bf16x3", 64),
                                           ("fp16x3", "fp16x3", 64),
                                           ("fp16x3", "bf16x3", 64), ("bf16", "bf16", 64)])
def test_planning_step_in_one_call(fwd, adj, rows):
    """pbh_forward_objective_adjoint (the objective formed in the forward's combine over
    every tile) against pbh_forward_shaped + pbh_objective_adjoint: the dose and
    dJ/dw bit for bit, J to 1e-12 (its FP64 sum is grouped by tile bands instead of by
    grid stride). Layer chunks of 2 (partial tiles on every slice with more than two
    active layers) and of all layers, with and without voxel weights, a grid whose x
    and y are not multiples of the tile (scalar edge path) and one that is, twice in a
    row on the same workspace."""
    from pbdose import engines, planning
    for nx, ny in ((100, 76), (128, 64)):
        print(f"case planning step {fwd}/{adj} {rows} rows {nx}x{ny}", flush=True)
        p = small_problem(nx=nx, ny=ny, nz=30, d=0.2, nsx=15, nsy=13, ds=0.5,
                          energies=tuple(np.linspace(75.0, 125.0, 7)))
        dp = engines.DeviceProblem(p)
        D = dp.forward()
        Dp, wv = planning.central_cube(p.grid, half=1.5, dose_value=float(D.max()) * 0.5)
        for chunk in (2, 40):
            D0 = dp.forward(fwd, layer_chunk=chunk, tile_rows=rows).clone()
            for vw in (wv, None):
                J0, g0 = planning.objective_gradient(dp, D0, Dp, vw, precision=adj)
                for _ in range(2):
                    D1, J1, g1 = planning.forward_objective_gradient(dp, Dp, vw, precision=adj, forward_precision=fwd,
                                                                     layer_chunk=chunk, tile_rows=rows)
                    assert torch.equal(D1, D0)
                    assert torch.equal(g1, g0)
                    assert abs(float(J1) / float(J0) - 1.0) <= 1e-12


def test_planning_step_rejects_bad_arguments():
    from pbdose import engines, planning
    p = small_problem()
    dp = engines.DeviceProblem(p)
    Dp, wv = planning.central_cube(p.grid)
    with pytest.raises(ValueError):         # no near-peak adjoint
        planning.forward_objective_gradient(dp, Dp, wv, precision="fp16x3_peak")
    with pytest.raises(ValueError):
        planning.forward_objective_gradient(dp, Dp[:1].contiguous(), wv)
    with pytest.raises(RuntimeError):       # 64-row tiles need the 2 : 5 pitch geometry
        planning.forward_objective_gradient(dp, Dp, wv, precision="bf16x3", tile_rows=64)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [((1, 1, 1), (1, 1), (120.0,)), ((48, 40, 1), (9, 7), (90.0, 120.0, 150.0)),
                                   ((1, 40, 6), (5, 9), (100.0, 130.0)), ((40, 1, 6), (9, 5), (100.0, 130.0))])
def test_adjoint_degenerate_shapes(shape, precision):
    """One voxel / spot / layer, one depth slice, one-voxel-wide grids: dJ/dw of a
    random residual vs the FP64 oracle adjoint."""
    from pbdose import engines, model, planning
    (nx, ny, nz), (nsx, nsy), E = shape
    print(f"case adjoint degenerate {shape[0]} {precision}", flush=True)
    p = small_problem(seed=6, nx=nx, ny=ny, nz=max(nz, 2), nsx=nsx, nsy=nsy, energies=E)
    if nz == 1:
        # the oracle's moments need two depths: tables of the first slice of two
        g = p.grid
        g1 = model.DoseGrid(g.nx, g.ny, 1, g.dx, g.dy, g.dz, g.ox, g.oy, g.oz)
        p = model.PBProblem(g1, p.lattice, p.idd[:, :1].copy(), p.sigma[:, :1].copy())
    dp = engines.DeviceProblem(p)
    R = np.random.default_rng(9).standard_normal(dp.shape).astype(np.float32)
    ref = O.adjoint_w(p.grid.x, p.grid.y, p.lattice.spot_x, p.lattice.spot_y, p.idd, p.sigma,
                      R.astype(np.float64))
    g = planning.adjoint(dp, torch.as_tensor(R, device=dp.dev), precision=precision)
    e = rel(g.cpu(), ref)
    print(f"  vs oracle {e:.2e}", flush=True)
    assert e <= (2e-5 if precision == "fp32" else MIXED_ADJ_TOL[precision])


def test_adjoint_split_autotune(tmp_path, monkeypatch):
    """DeviceProblem.autotune_split times the dispatch model's split and its
    neighbours; later automatic adjoints use the winner (dJ/dw equals the explicit
    split's bit for bit, as every split does), and the cache returns it without
    timing again."""
    from pbdose import engines, planning
    monkeypatch.setenv("PBDOSE_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    print("case adjoint split autotune bf16x3", flush=True)
    p = small_problem(nx=64, ny=48, nz=40, d=0.2, nsx=13, nsy=11, ds=0.5,
                      energies=tuple(np.linspace(70.0, 130.0, 8)))
    dp = engines.DeviceProblem(p)
    D = dp.forward()
    Dp, wv = planning.central_cube(p.grid, half=1.5, dose_value=float(D.max()) * 0.5)
    best = dp.autotune_split("bf16x3", reps=3)
    assert best in [sp for _, sp in dp.split_results] and 1 <= best <= 16
    J0, g0 = planning.objective_gradient(dp, D, Dp, wv, precision="bf16x3")
    J1, g1 = planning.objective_gradient(dp, D, Dp, wv, precision="bf16x3", split=best)
    assert torch.equal(g0, g1) and float(J0) == float(J1)
    dp2 = engines.DeviceProblem(p)
    assert dp2.autotune_split("bf16x3") == best and not hasattr(dp2, "split_results")
