"""Robust-scenario batches (GPU, through pbh_batch_*): every (field, scenario)
item against the single-problem entry points, the FP64 oracle on the modified
problem that expresses the scenario, and HIP-graph replays against eager runs.

Bars: FP32 dose <= 1e-6 and bf16x3 dose <= 2e-5 max global relative error vs
the FP64 oracle; dJ/dw <= 2e-5 (relative to its largest magnitude) vs the
oracle adjoint; batch vs single-problem path <= 1e-6 (the same kernels; only the
zero padding of narrower spot chunks can differ); graph replays bit-identical
to eager runs.
"""
import numpy as np
import pytest
import torch

from oracle import pb_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_fields():
    from pbdose import model, robust
    g = model.DoseGrid(48, 40, 24, 0.25, 0.25, 0.25, -(47) * 0.25 / 2, -(39) * 0.25 / 2, 0.0)
    fa = robust.Field(g, model.SpotLattice(9, 7, 3, 0.5, 0.5, energies=np.array([80.0, 100.0, 120.0]),
                                           weights=np.random.default_rng(0).lognormal(0, 0.5, (3, 7, 9))))
    # 0.37 cm pitch: no small voxel/spot ratio, factors evaluated directly
    fb = robust.Field(g, model.SpotLattice(6, 11, 2, 0.37, 0.37, 0.1, -0.2, energies=np.array([90.0, 115.0]),
                                           weights=np.random.default_rng(1).lognormal(0, 0.5, (2, 11, 6))))
    return [fa, fb]


def fields_with_25_pitch():
    """small_fields plus a field on 2 mm voxels / 5 mm spots, the geometry of the
    64-row mixed tiles."""
    from pbdose import model, robust
    g = model.DoseGrid(72, 70, 20, 0.2, 0.2, 0.2, -(71) * 0.2 / 2, -(69) * 0.2 / 2, 0.0)
    fc = robust.Field(g, model.SpotLattice(13, 11, 3, 0.5, 0.5, 0.2, 0.1, energies=np.array([85.0, 110.0, 135.0]),
                                           weights=np.random.default_rng(2).lognormal(0, 0.5, (3, 11, 13))))
    return small_fields() + [fc]


def small_scenarios():
    from pbdose import robust
    S = robust.Scenario
    return [S(), S(0.3, 0, 0, 1.0), S(0, -0.3, 0, 0.965), S(0, 0, 0.3, 1.035), S(0, 0, -0.3, 1.0)]


def oracle_item(field, sc, z_idx=None):
    """FP64 oracle dose and layer tables of one field in one scenario."""
    g, lat = field.grid, field.lattice
    z = g.oz + sc.dz + np.arange(g.nz) * g.dz
    E = np.asarray(lat.energies, dtype=np.float64)
    if sc.range_scale != 1.0:
        E = O.energy_from_range(O.csda_range(E) * sc.range_scale)
    idd, sig = O.layer_luts(z, E)
    sx = O.spot_axis(lat.n_sx, lat.dx_spot, lat.x0 + sc.dx)
    sy = O.spot_axis(lat.n_sy, lat.dy_spot, lat.y0 + sc.dy)
    w = lat.weights.astype(np.float32).astype(np.float64)
    D = O.separable_dose(g.x, g.y, sx, sy, w, idd, sig, z_idx=z_idx)
    return D, (sx, sy, idd, sig)


def adjoint_form(precision):
    """The adjoint arithmetic a batch pairs with a forward form: the forward's 16-bit
    type, single products for the plain forms, split products otherwise."""
    if precision == "fp32":
        return "fp32"
    t = "fp16" if precision.startswith("fp16") else "bf16"
    return t if precision in ("bf16", "fp16") else t + "x3"


def single_path(b, k, precision):
    """Dose, J and dJ/dw of item k through pbh_forward_tuned / pbh_objective / pbh_adjoint_tuned."""
    from pbdose import engines, planning
    fi, _ = b.items[k]
    dp = engines.DeviceProblem(b.problems[k], b.dev, weights=b.w[fi], idd=b.idd[k], sigma=b.sigma[k])
    D = dp.forward(precision, layer_chunk=b.tuning["fwd_chunk"])
    J, R = planning.objective(dp, D, b.Dp[fi], b.wv[fi])
    try:
        g = planning.adjoint(dp, R, precision=adjoint_form(precision))
    except RuntimeError:
        # no table geometry (direct factors): the batch uses the FP32 adjoint for this item
        g = planning.adjoint(dp, R)
    return D, J, g


@pytest.mark.parametrize("precision,tuning", [
    ("fp32", {"fwd_chunk": 0, "adj_split": 0, "group": 1, "order": 1}),
    ("fp32", {"fwd_chunk": 2, "adj_split": 2, "group": 0, "order": 0}),
    ("bf16x3", {"fwd_chunk": 4096, "adj_split": 1, "group": 0, "order": 1}),
    ("fp16x3", {"fwd_chunk": 2, "adj_split": 0, "group": 1, "order": 0}),
    ("bf16x3_peak", {"fwd_chunk": 0, "adj_split": 0, "group": 0, "order": 1}),
])
def test_small_batch_matches_oracle_and_single_path(precision, tuning):
    from pbdose import robust
    print(f"case small robust batch {precision} {tuning}", flush=True)
    fields, scen = small_fields(), small_scenarios()
    b = robust.RobustBatch(fields, scen, precision=precision, tuning=tuning)
    b.run_eager()
    torch.cuda.synchronize()
    check_items(b, fields, scen, precision)


def check_items(b, fields, scen, precision):
    """Every item of a run batch vs the FP64 oracle (dose, dJ/dw, J) and vs the
    single-problem path."""
    # the dose bounds of the single-problem forms (tests/test_gpu_forward.py)
    tol = {"fp32": 1e-6, "bf16x3": 2e-5, "fp16x3": 2e-6, "bf16x3_peak": 5e-3, "bf16": 5e-3,
           "fp16": 1e-3}[precision]
    worst = [0.0] * 5
    for k, (fi, si) in enumerate(b.items):
        D = b.dose[k].double().cpu().numpy()
        ref, (sx, sy, idd, sig) = oracle_item(fields[fi], scen[si])
        e_d = O.relative_error(D, ref)["max_rel_err_global"]
        Dp, wv = b.Dp[fi].double().cpu().numpy(), b.wv[fi].double().cpu().numpy()
        Rf = (2.0 * wv * (D - Dp)).astype(np.float32).astype(np.float64)
        g_ref = O.adjoint_w(fields[fi].grid.x, fields[fi].grid.y, sx, sy, idd, sig, Rf)
        e_g = rel(b.grad_w[k].cpu(), g_ref)
        e_j = abs(float(b.J[k]) - O.objective(D, Dp, wv)) / O.objective(D, Dp, wv)
        D1, J1, g1 = single_path(b, k, precision)
        e_s = rel(b.dose[k].cpu(), D1.cpu())
        e_sg = rel(b.grad_w[k].cpu(), g1.cpu())
        worst = [max(x, y) for x, y in zip(worst, (e_d, e_g, e_j, e_s, e_sg))]
        assert e_d <= tol, (k, e_d)
        assert e_g <= {"bf16": 1e-2, "fp16": 3e-3}.get(adjoint_form(precision), 2e-5), (k, e_g)
        assert e_j <= 1e-9, (k, e_j)
        assert e_s <= 1e-6 and e_sg <= 1e-6, (k, e_s, e_sg)
        assert abs(float(b.J[k]) - float(J1)) <= 1e-6 * abs(float(J1))
    print(f"  {len(b.items)} items, {b.launch_count} launches: dose vs oracle {worst[0]:.2e}, "
          f"dJ/dw vs oracle {worst[1]:.2e}, J {worst[2]:.1e}, vs single path {worst[3]:.1e} / {worst[4]:.1e}",
          flush=True)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_degenerate_fields_in_a_batch(precision):
    """Fields on a 1 x 1 x 2 grid (one spot, one layer), a one-voxel-wide column and a
    one-voxel-wide row, at 2.5 mm voxels and 5 mm spots, in one batch with two scenarios."""
    from pbdose import model, robust
    print(f"case degenerate fields in a batch {precision}", flush=True)
    rng = np.random.default_rng(5)

    def field(nx, ny, nz, nsx, nsy, E):
        g = model.DoseGrid(nx, ny, nz, 0.25, 0.25, 0.25, -(nx - 1) * 0.25 / 2, -(ny - 1) * 0.25 / 2, 0.0)
        w = rng.lognormal(0.0, 0.5, (len(E), nsy, nsx))
        return robust.Field(g, model.SpotLattice(nsx, nsy, len(E), 0.5, 0.5, energies=np.array(E), weights=w))

    fields = [field(1, 1, 2, 1, 1, [120.0]), field(1, 40, 6, 5, 9, [100.0, 130.0]),
              field(40, 1, 6, 9, 5, [100.0, 130.0])]
    scen = small_scenarios()[:2]
    b = robust.RobustBatch(fields, scen, precision=precision, tuning=robust.DEFAULT_TUNING)
    b.run_eager()
    torch.cuda.synchronize()
    check_items(b, fields, scen, precision)


def test_forward_only_batch_and_lut_gradients():
    from pbdose import engines, planning, robust
    print("case forward-only batch; batch with dJ/dIDD, dJ/dsigma", flush=True)
    fields, scen = small_fields(), small_scenarios()[:3]
    b = robust.RobustBatch(fields, scen, precision="fp32", objective=False, tuning=robust.DEFAULT_TUNING)
    b.run_eager()
    torch.cuda.synchronize()
    assert 1 <= b.launch_count <= 2      # forward launches only: table and direct-factor groups
    for k, (fi, si) in enumerate(b.items):
        ref, _ = oracle_item(fields[fi], scen[si])
        assert O.relative_error(b.dose[k].double().cpu().numpy(), ref)["max_rel_err_global"] <= 1e-6
    bl = robust.RobustBatch(fields, scen, precision="fp32", lut_grads=True, tuning=robust.DEFAULT_TUNING)
    bl.run_eager()
    torch.cuda.synchronize()
    for k, (fi, si) in enumerate(bl.items):
        dp = engines.DeviceProblem(bl.problems[k], bl.dev, weights=bl.w[fi], idd=bl.idd[k], sigma=bl.sigma[k])
        _, R = planning.objective(dp, bl.dose[k], bl.Dp[fi], bl.wv[fi])
        gw, gi, gs = planning.adjoint(dp, R, lut_grads=True)
        assert rel(bl.grad_idd[k].cpu(), gi.cpu()) <= 1e-12
        assert rel(bl.grad_sigma[k].cpu(), gs.cpu()) <= 1e-12
        assert rel(bl.grad_w[k].cpu(), gw.cpu()) <= 1e-6


@pytest.mark.parametrize("variant", [{}, {"tile_rows": 64, "persistent": 1, "fwd_chunk": 2}])
def test_graph_replay_equals_eager_runs(variant):
    """Graph replays equal eager runs bit for bit, also with the persistent forward
    launch, whose claim counters every launch leaves at zero for the next."""
    from pbdose import robust
    print(f"case robust batch graph replays {variant}", flush=True)
    fields, scen = fields_with_25_pitch(), small_scenarios()
    b = robust.RobustBatch(fields, scen, precision="bf16x3", tuning={**robust.DEFAULT_TUNING, **variant})
    b.capture()
    for rep in range(3):
        # change the inputs between replays: the graph must read them live
        for w in b.w:
            w.mul_(1.0 + 0.25 * rep)
        b.run_eager()
        torch.cuda.synchronize()
        eager = ([d.clone() for d in b.dose], [g.clone() for g in b.grad_w], b.J.clone())
        for d in b.dose:
            d.fill_(-1.0)
        for g in b.grad_w:
            g.fill_(-1.0)
        b.J.fill_(-1.0)
        b.run()
        torch.cuda.synchronize()
        assert all(torch.equal(x, y) for x, y in zip(b.dose, eager[0])), rep
        assert all(torch.equal(x, y) for x, y in zip(b.grad_w, eager[1])), rep
        assert torch.equal(b.J, eager[2]), rep
    print(f"  3 replays bit-identical to eager runs ({b.launch_count} kernel nodes)", flush=True)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "bf16"])
def test_tile_height_and_persistent_launch_change_no_bit(precision):
    """The forward of a batch with 64-row tiles and/or a persistent launch (work
    units claimed from a counter) equals the 32-row, one-workgroup-per-unit launch
    bit for bit, at the same layer chunk; each is checked against the oracle."""
    from pbdose import robust
    print(f"case batch tile heights / persistent launch {precision}", flush=True)
    fields, scen = fields_with_25_pitch(), small_scenarios()
    ref = None
    for rows in (32, 64):
        for pers in (0, 1):
            t = {**robust.DEFAULT_TUNING, "fwd_chunk": 2, "tile_rows": rows, "persistent": pers}
            b = robust.RobustBatch(fields, scen, precision=precision, tuning=t)
            b.run_eager()
            b.run_eager()            # a second launch starts from the counters the first reset
            torch.cuda.synchronize()
            dose = [d.clone() for d in b.dose]
            if ref is None:
                ref = dose
                check_items(b, fields, scen, precision)
            else:
                assert all(torch.equal(x, y) for x, y in zip(dose, ref)), (rows, pers)
            b.close()


@pytest.mark.skip(reason="the objective formed at the batch forward's tile store was taken out: a GPU run "
                         "of the bf16x3 / 64-row case wrote nothing for 180 s and was stopped, and no cause "
                         "was found by reading the code; batches keep the objective kernel")
@pytest.mark.parametrize("precision,rows", [("fp32", 32), ("bf16x3", 64), ("bf16x3", 32)])
def test_objective_at_the_tile_store(precision, rows):
    """Items forwarded in one layer chunk would form R and J's partial sums where the
    forward stores its dose tiles: no objective launch, J summed in the depth
    reduction's launch. Every item against the FP64 oracle and the single-problem path
    (which runs the objective kernel), with and without a persistent launch, on two runs
    in a row."""
    from pbdose import robust
    print(f"case objective at the tile store {precision} {rows} rows", flush=True)
    fields, scen = fields_with_25_pitch(), small_scenarios()
    for pers in (0, 1):
        t = {**robust.DEFAULT_TUNING, "fwd_chunk": 4096, "tile_rows": rows, "persistent": pers}
        b = robust.RobustBatch(fields, scen, precision=precision, tuning=t)
        b0 = robust.RobustBatch(fields, scen, precision=precision, tuning={**t, "fwd_chunk": 1})
        # against one layer per chunk: no combine, objective and J-sum launches
        assert b.launch_count == b0.launch_count - 3, (b.launch_count, b0.launch_count)
        b0.close()
        for _ in range(2):
            b.run_eager()
        torch.cuda.synchronize()
        check_items(b, fields, scen, precision)
        b.close()


def test_autotune_picks_a_measured_candidate(tmp_path, monkeypatch):
    from pbdose import robust
    monkeypatch.setenv("PBDOSE_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    fields, scen = small_fields(), small_scenarios()[:2]
    cands = [{"fwd_chunk": 0, "adj_split": 1, "group": 1, "order": 0},
             {"fwd_chunk": 1, "adj_split": 4, "group": 0, "order": 1}]
    b = robust.RobustBatch(fields, scen, precision="fp32", tuning=robust.DEFAULT_TUNING)
    best = b.autotune(candidates=cands, reps=2)
    assert best in cands
    assert (tmp_path / "tune.json").exists()
    assert b.autotune(candidates=cands[:1]) == best          # served from the cache


def test_robust_workload_128cube_63_items():
    """3 fields (31x31x25, 41x41x40, 27x35x45) x 21 scenarios on the 128^3 grid,
    bf16x3 forward + FP32 adjoint: every item vs the single-problem path, and
    three items vs the FP64 oracle on every 16th slice."""
    from pbdose import robust
    print("case robust workload 63 items on 128^3", flush=True)
    fields, scen = robust.robust_fields(), robust.standard_scenarios()
    b = robust.RobustBatch(fields, scen, precision="bf16x3", tuning=robust.DEFAULT_TUNING)
    b.run()
    torch.cuda.synchronize()
    assert len(b.items) == 63
    ws, wg = 0.0, 0.0
    for k in range(63):
        D1, J1, g1 = single_path(b, k, "bf16x3")
        ws = max(ws, rel(b.dose[k].cpu(), D1.cpu()))
        wg = max(wg, rel(b.grad_w[k].cpu(), g1.cpu()))
        assert abs(float(b.J[k]) - float(J1)) <= 1e-6 * abs(float(J1))
    print(f"  batch vs single path: dose {ws:.1e}, dJ/dw {wg:.1e}", flush=True)
    assert ws <= 1e-6 and wg <= 1e-6
    # one persistent launch of 64-row tiles over the same 63 items: the same doses
    t = {**robust.DEFAULT_TUNING, "fwd_chunk": b.tuning["fwd_chunk"] or 4096, "group": 0, "tile_rows": 64,
         "persistent": 1}
    b2 = robust.RobustBatch(fields, scen, precision="bf16x3", tuning=t)
    b2.run()
    torch.cuda.synchronize()
    b3 = robust.RobustBatch(fields, scen, precision="bf16x3", tuning={**t, "tile_rows": 32, "persistent": 0})
    b3.run()
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(b2.dose, b3.dose))
    print(f"  persistent 64-row forward: {b2.launch_count} launches, doses equal the 32-row launch", flush=True)
    b2.close()
    b3.close()
    zs = np.arange(0, 128, 16)
    for k in (1, 30, 62):
        fi, si = b.items[k]
        ref, _ = oracle_item(fields[fi], scen[si], z_idx=zs)
        e = O.relative_error(b.dose[k].double().cpu().numpy()[zs], ref)["max_rel_err_global"]
        print(f"  item {k} (field {fi}, {scen[si].name}): vs FP64 oracle {e:.2e}", flush=True)
        assert e <= 2e-5
