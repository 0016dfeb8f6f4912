"""The C-ABI library: it loads, exports every declared symbol, and its pure
host entry points (error strings, validation, workspace sizing) behave.
No kernel is launched here."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from pbdose import _lib

HEADER = os.path.join(ROOT, "include", "pbdose_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in pbdose_hip.h but not exported"
    # the ctypes table covers exactly the header
    assert set(names) == set(_lib.SIGNATURES)


def test_error_strings():
    assert _lib.strerror(0) == "ok"
    assert "null" in _lib.strerror(-1)
    assert "workspace" in _lib.strerror(-7)
    assert _lib.strerror(-99) == "unknown error"
    assert b"gfx950" in _lib.lib().pbh_version()


def test_validate_rejects_bad_problems():
    L = _lib.lib()
    g = _lib.PbhGrid(128, 128, 128, 0, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = _lib.PbhLattice(41, 41, 40, 0, 0.5, 0.5, 0.0, 0.0)
    assert L.pbh_validate(g, lat) == 0
    assert L.pbh_validate(None, lat) == -1
    bad = _lib.PbhGrid(0, 128, 128, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    assert L.pbh_validate(bad, lat) == -2
    bad = _lib.PbhGrid(128, 128, 128, 0, -0.2, 0.2, 0.2, 0, 0, 0)
    assert L.pbh_validate(bad, lat) == -2
    badl = _lib.PbhLattice(41, 41, 40, 0, 0.0, 0.5, 0.0, 0.0)
    assert L.pbh_validate(g, badl) == -2
    huge = _lib.PbhGrid(100000, 128, 128, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    assert L.pbh_validate(huge, lat) == -5


def test_forward_rejects_null_and_bad_precision_without_launching():
    L = _lib.lib()
    g = _lib.PbhGrid(8, 8, 8, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    lat = _lib.PbhLattice(3, 3, 2, 0, 0.5, 0.5, 0.0, 0.0)
    nul = C.c_void_p(0)
    assert L.pbh_forward(g, lat, nul, nul, nul, nul, 0, nul, 0, nul) == -1
    one = C.c_void_p(16)   # never dereferenced: the precision check fails first
    assert L.pbh_forward(g, lat, one, one, one, one, 99, nul, 0, nul) == -5
    # the profiled forward needs both events; the no-workspace forward checks its
    # arguments and its scratch limit before touching the device
    assert L.pbh_forward_profiled(g, lat, one, one, one, one, 3, 0, 0, nul, 0, nul, one, nul) == -1
    assert L.pbh_forward_noalloc(g, lat, nul, nul, nul, nul, 3, 0, nul) == -1
    assert L.pbh_forward_noalloc(g, lat, one, one, one, one, 99, 0, nul) == -5
    big = _lib.PbhGrid(512, 512, 512, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    blat = _lib.PbhLattice(41, 41, 40, 0, 0.5, 0.5, 0.0, 0.0)
    assert L.pbh_forward_workspace_bytes_tuned(big, blat, 3, 1) > _lib.PBH_NOALLOC_SCRATCH_BYTES
    assert L.pbh_forward_noalloc(big, blat, one, one, one, one, 3, 1, nul) == -8
    assert L.pbh_forward_fp64(g, lat, one, nul, one, one, nul) == -1
    empty = _lib.PbhGrid(0, 8, 8, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    assert L.pbh_forward_fp64(empty, lat, one, one, one, one, nul) == -2


def test_average_sigma_rejects_bad_arguments_without_launching():
    L = _lib.lib()
    nul, one = C.c_void_p(0), C.c_void_p(16)
    # stand-in pointers, never dereferenced: each call below fails a check first
    args = [one, one, one, 3, 16, 9, 9, one, one, one, one, nul]
    for f in (L.pbh_average_sigma, L.pbh_average_sigma_fp64):
        bad = list(args); bad[0] = nul
        assert f(*bad) == -1
        bad = list(args); bad[7] = nul
        assert f(*bad) == -1
        bad = list(args); bad[10] = nul
        assert f(*bad) == -1
        for k, v in ((3, 0), (4, 0), (5, -1), (6, 0), (4, 100000), (6, 5000)):
            bad = list(args); bad[k] = v
            assert f(*bad) == -2, (k, v)
        bad = list(args); bad[4], bad[5], bad[6] = 4096, 1024, 1024      # nz x spots > 2^31
        assert f(*bad) == -5


def test_forward_workspace_query():
    L = _lib.lib()
    g = _lib.PbhGrid(128, 128, 128, 0, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = _lib.PbhLattice(41, 41, 40, 0, 0.5, 0.5, 0.0, 0.0)
    one = L.pbh_forward_workspace_bytes_tuned(g, lat, 3, 40)
    assert one >= 40 * 2 * 48 * 64 * 2                       # split 16-bit weights
    # chunks of 10 layers: 3 partial tiles of 32 x 128 FP32 per dose tile
    assert L.pbh_forward_workspace_bytes_tuned(g, lat, 3, 10) - one >= 3 * 512 * 32 * 128 * 4
    assert L.pbh_forward_workspace_bytes_tuned(g, lat, 3, 4096) == one
    assert L.pbh_forward_workspace_bytes_tuned(g, lat, 3, -1) == 0
    assert L.pbh_forward_workspace_bytes(g, lat, 3) >= one


def test_lut_workspace_query():
    from pbdose import physics
    L = _lib.lib()
    m = physics.beam_model()
    n = L.pbh_lut_workspace_bytes(m, 128, 0.0, 0.2, 40)
    assert n > 40 * 6000 * 16   # binning + Gaussian kernel per layer
    # the Fermi-Eyges fine grid follows the depth span: a coarse grid needs more
    assert L.pbh_lut_workspace_bytes(m, 128, 0.0, 3.0, 40) > n
    # one depth: the moment grid is [0, z0] (physics.py:247-249)
    assert L.pbh_lut_workspace_bytes(m, 1, 12.3, 1.0, 40) > 0
    assert L.pbh_lut_workspace_bytes(m, 0, 0.0, 0.2, 40) == 0
    assert L.pbh_lut_workspace_bytes(m, 128, 0.0, 0.0, 40) == 0
    # the tangent build holds a derivative beside every table entry
    assert L.pbh_lut_tangent_workspace_bytes(m, 128, 0.0, 0.2, 40) > n
    bad = physics.beam_model()
    bad.n_fine_idd = 1
    assert L.pbh_lut_workspace_bytes(bad, 128, 0.0, 0.2, 40) == 0
    assert L.pbh_build_lut(m, 128, 0.0, 0.2, 40, None, None, None, None, None, None, None, 0,
                           None) == -1
    assert L.pbh_build_lut_tangent(m, 128, 0.0, 0.2, 40, None, None, None, None, None, None, None,
                                   None, None, None, 0, None) == -1
    assert L.pbh_interp_table_slope(None, None, 10, None, None, None, 1, None) == -1
    # depth-dependent straggling: extra cumulative-integral and amplitude tables
    d = physics.beam_model(physics.IDDModel(depth_straggling=True))
    assert d.straggling_depth == 1
    assert L.pbh_lut_workspace_bytes(d, 128, 0.0, 0.2, 40) >= n + 40 * (4000 + 6000) * 8
    icru = physics.beam_model(physics.IDDModel(straggling_model="icru49"))
    icru.straggling_depth = 1                  # defined for the Bohr integral only
    assert L.pbh_lut_workspace_bytes(icru, 128, 0.0, 0.2, 40) == 0
    d.straggling_depth = 2
    assert L.pbh_lut_workspace_bytes(d, 128, 0.0, 0.2, 40) == 0


def test_product_refuses_cpu_device():
    import numpy as np
    import torch
    from pbdose import engines, model
    g = model.DoseGrid(4, 4, 4, 0.2, 0.2, 0.2)
    lat = model.SpotLattice(2, 2, 1, 0.5, 0.5, weights=np.ones((1, 2, 2)))
    prob = model.PBProblem(g, lat, np.ones((1, 4)), np.ones((1, 4)))
    with pytest.raises(ValueError, match="GPU only"):
        engines.torch_separable(prob, "cpu")
    # options the reference accepts that have no effect here warn before anything runs
    for kw, text in (({"layer_chunk": 4}, "layer_chunk"), ({"allow_tf32": True}, "allow_tf32"),
                     ({"return_intermediates": True}, "return_intermediates")):
        with pytest.warns(UserWarning, match=text), pytest.raises(ValueError, match="GPU only"):
            engines.torch_separable(prob, "cpu", **kw)
    # float64 is the FP64 form; other dtypes and the float32 arithmetic forms refuse
    for kw, text in (({"dtype": torch.float16}, "not supported"),
                     ({"dtype": torch.float64, "precision": "bf16x3"}, "FP64")):
        with pytest.raises(ValueError, match=text):
            engines.torch_separable(prob, "cuda", **kw)


def _batch_items(n=3):
    """Host descriptors with stand-in pointers (never dereferenced: no kernel runs)."""
    arr = (_lib.PbhBatchItem * n)()
    for k in range(n):
        it = arr[k]
        it.grid = _lib.PbhGrid(64, 48, 32, 0, 0.2, 0.2, 0.2, -6.3, -4.7, 0.0)
        it.lattice = _lib.PbhLattice(9 + k, 7, 2 + k, 0, 0.5, 0.5, 0.3 * k, 0.0)
        it.w = it.idd = it.sigma = it.dose = 256
        it.presc = 256 if k != 1 else None
        it.grad_w = 256
    return arr


def test_batch_workspace_query_and_argument_checks():
    L = _lib.lib()
    items = _batch_items()
    t = _lib.PbhBatchTuning(4096, 0, 1, 1)                              # one layer chunk
    n = L.pbh_batch_workspace_bytes(items, 3, 3, C.byref(t))
    # at least: the residual and per-slice partials of the two items with a prescription
    need = sum(64 * 48 * 32 * 4 + 32 * (2 + k) * 7 * (9 + k) * 4 for k in (0, 2))
    assert n >= need
    assert L.pbh_batch_workspace_bytes(items, 3, 3, None) > 0          # default tuning
    assert L.pbh_batch_workspace_bytes(items, 0, 3, None) == 0          # empty batch
    assert L.pbh_batch_workspace_bytes(items, 3, 99, None) == 0         # bad precision
    bad = _lib.PbhBatchTuning(0, 17, 0, 0)                              # adjoint split above 16
    assert L.pbh_batch_workspace_bytes(items, 3, 3, C.byref(bad)) == 0
    bad = _lib.PbhBatchTuning(-1, 0, 0, 0)                              # negative layer chunk
    assert L.pbh_batch_workspace_bytes(items, 3, 3, C.byref(bad)) == 0
    # layer chunks of one layer need room for the partial tiles of later chunks
    t1 = _lib.PbhBatchTuning(1, 0, 1, 1)
    assert L.pbh_batch_workspace_bytes(items, 3, 3, C.byref(t1)) > n
    items[1].dose = None
    assert L.pbh_batch_workspace_bytes(items, 3, 3, None) == 0
    items[1].dose = 256
    items[0].grad_w = None                                              # objective without a gradient
    assert L.pbh_batch_workspace_bytes(items, 3, 3, None) == 0


def test_batch_create_rejects_before_touching_the_device():
    L = _lib.lib()
    items = _batch_items()
    out = C.c_void_p()
    t = _lib.PbhBatchTuning(4096, 0, 1, 1)
    assert L.pbh_batch_create(items, 3, 3, C.byref(t), None, 0, None, C.byref(out)) == -1
    assert L.pbh_batch_create(items, 3, 3, C.byref(t), C.c_void_p(256), 16, None, C.byref(out)) == -7
    assert L.pbh_batch_create(items, 3, 3, C.byref(t), C.c_void_p(256), 16, None, None) == -1
    assert not out.value
    assert L.pbh_batch_forward_adjoint(None, None) == -1
    assert L.pbh_batch_launch_count(None) == -1
    L.pbh_batch_destroy(None)


def test_standard_scenarios_and_fields():
    from pbdose import robust
    sc = robust.standard_scenarios()
    assert len(sc) == 21 and len(set(sc)) == 21
    assert robust.Scenario() in sc
    assert sorted({s.range_scale for s in sc}) == [0.965, 1.0, 1.035]
    assert sum(1 for s in sc if (s.dx, s.dy, s.dz) == (0, 0, 0)) == 3
    assert all(abs(s.dx) + abs(s.dy) + abs(s.dz) in (0.0, 0.3) for s in sc)
    f = robust.robust_fields()
    assert [(x.lattice.n_sx, x.lattice.n_sy, x.lattice.n_layers) for x in f] == [(31, 31, 25), (41, 41, 40), (27, 35, 45)]
    assert sum(x.lattice.n_spots for x in f) == 24025 + 67240 + 42525
    assert f[2].lattice.weights.shape == (45, 35, 27)


@pytest.mark.parametrize("spacing,dta_mm", [((0.2, 0.2, 0.2), 2.0), ((0.2, 0.2, 0.2), 3.0),
                                            ((0.25, 0.2, 0.15), 3.0), ((0.1, 0.1, 0.1), 10.0),
                                            ((0.3, 0.3, 0.3), 2.0), ((0.2, 0.2, 0.2), 1.0)])
def test_gamma_offset_count_matches_oracle(spacing, dta_mm):
    from oracle import pb_oracle as O
    dz, dy, dx = spacing
    g = _lib.PbhGrid(8, 8, 8, 0, dx, dy, dz, 0.0, 0.0, 0.0)
    assert _lib.lib().pbh_gamma_count_offsets(g, dta_mm) == len(O.dta_offsets(spacing, dta_mm / 10.0)[0])


def test_gamma_argument_checks():
    L = _lib.lib()
    g = _lib.PbhGrid(8, 8, 8, 0, 0.2, 0.2, 0.2, 0.0, 0.0, 0.0)
    assert L.pbh_gamma_workspace_bytes(g) > 0
    nul = C.c_void_p(0)
    one = C.c_void_p(256)
    assert L.pbh_gamma_index(g, nul, 0, one, 0, 2.0, 2.0, 10.0, 0.0, 0, nul, one, one, one, 1 << 20, nul) == -1
    assert L.pbh_gamma_index(g, one, 2, one, 0, 2.0, 2.0, 10.0, 0.0, 0, nul, one, one, one, 1 << 20, nul) == -5
    assert L.pbh_gamma_index(g, one, 0, one, 0, 0.0, 2.0, 10.0, 0.0, 0, nul, one, one, one, 1 << 20, nul) == -2
    assert L.pbh_gamma_index(g, one, 0, one, 0, 2.0, 2.0, 10.0, 0.0, 0, nul, one, one, one, 16, nul) == -7


def test_weight_step_argument_checks():
    L = _lib.lib()
    nul = C.c_void_p(0)
    one = C.c_void_p(256)
    assert L.pbh_weight_step(nul, one, one, one, 2, 10, 0, 0.1, 0.0, 1.0, nul) == -1
    assert L.pbh_weight_step(one, one, one, nul, 2, 10, 1, 0.1, 0.0, 1.0, nul) == -1   # worst needs J
    assert L.pbh_weight_step(one, one, one, one, 0, 10, 0, 0.1, 0.0, 1.0, nul) == -2
    assert L.pbh_weight_step(one, one, one, one, 2, 10, 7, 0.1, 0.0, 1.0, nul) == -5
    assert L.pbh_weight_step(one, one, one, one, 2, 10, 0, 0.1, 2.0, 1.0, nul) == -2   # w_min > w_max


def test_forward_chunk_chooser():
    """Host-only layer-chunk model: argument checks, no chunking when every slice
    has all layers active and the tiles fill the resident slots exactly, and the
    chunks measured fastest on the 128^3 / 41x41 x 40-layer field."""
    import numpy as np
    from oracle import pb_oracle as O
    from pbdose import model
    L = _lib.lib()
    g = model.DoseGrid(128, 128, 128, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5)
    gs, ls = _lib.grid_struct(g), _lib.lattice_struct(lat)
    full = np.full(128, 40, np.int32)
    assert L.pbh_forward_choose_chunk(gs, ls, 0, None, 256) == -1
    assert L.pbh_forward_choose_chunk(gs, ls, 9, full.ctypes.data, 256) == -5
    # 512 tiles on 2 x 256 slots: one round without chunking
    assert L.pbh_forward_choose_chunk(gs, ls, 0, full.ctypes.data, 256) == 40
    idd, _ = O.layer_luts(g.z, np.linspace(70, 180, 40))
    act = np.ascontiguousarray((idd != 0).sum(0).astype(np.int32))
    assert L.pbh_forward_choose_chunk(gs, ls, 0, act.ctypes.data, 256) == 26     # FP32
    for prec in (1, 3):                                                          # BF16, BF16x3
        assert 16 <= L.pbh_forward_choose_chunk(gs, ls, prec, act.ctypes.data, 256) <= 20


def test_objective_and_adjoint_reject_bad_arguments_without_launching():
    """Argument checks of pbh_objective and pbh_adjoint_tuned return their codes
    before any device call (the pointers below are never dereferenced)."""
    L = _lib.lib()
    g = _lib.PbhGrid(128, 128, 128, 0, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = _lib.PbhLattice(41, 41, 40, 0, 0.5, 0.5, 0.0, 0.0)
    nul, one = C.c_void_p(0), C.c_void_p(256)
    ob = L.pbh_objective_workspace_bytes(g)
    assert ob >= 1024 * 8                                  # FP64 partial sums of J
    assert L.pbh_objective(g, nul, one, nul, one, one, one, ob, nul) == -1          # no dose
    assert L.pbh_objective(g, one, one, nul, one, nul, one, ob, nul) == -1          # no J output
    bad = _lib.PbhGrid(0, 128, 128, 0, 0.2, 0.2, 0.2, 0, 0, 0)
    assert L.pbh_objective(bad, one, one, nul, one, one, one, ob, nul) == -2
    assert L.pbh_objective(g, one, one, nul, one, one, one, ob - 1, nul) == -7
    P = _lib.PREC
    ab = L.pbh_adjoint_workspace_bytes(g, lat, P["bf16x3"])
    assert ab >= 128 * 40 * 41 * 41 * 4                    # per-slice FP32 partials
    args = lambda prec, gi, split, nbytes: (g, lat, one, one, one, one, one, gi, nul, prec, split, one, nbytes, nul)
    assert L.pbh_adjoint_tuned(g, lat, nul, one, one, one, one, nul, nul, 0, 0, one, ab, nul) == -1   # no weights
    assert L.pbh_adjoint_tuned(*args(99, nul, 0, ab)) == -5                        # unknown precision
    assert L.pbh_adjoint_tuned(*args(P["fp16x3_peak"], nul, 0, ab)) == -5          # no near-peak adjoint
    assert L.pbh_adjoint_tuned(*args(P["bf16x3"], one, 0, ab)) == -5               # table gradients: FP32 only
    assert L.pbh_adjoint_tuned(*args(P["fp16x3"], one, 0, ab)) == -5
    assert L.pbh_adjoint_tuned(*args(P["bf16x3"], nul, -1, ab)) == -2              # negative split
    assert L.pbh_adjoint_tuned(*args(P["bf16x3"], nul, 0, ab - 1)) == -7           # workspace too small


def test_bench_runs_on_one_gpu_only():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says why,
    before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_adjoint_split_chooser():
    """Host-only layer-split model of the adjoint: argument checks, and the splits
    measured fastest on the 128^3 / 41x41 x 40-layer field (BF16x3 7, BF16 10,
    FP32 5 of splits 2-16)."""
    import numpy as np
    from oracle import pb_oracle as O
    from pbdose import model
    L = _lib.lib()
    g = model.DoseGrid(128, 128, 128, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5)
    gs, ls = _lib.grid_struct(g), _lib.lattice_struct(lat)
    idd, _ = O.layer_luts(g.z, np.linspace(70, 180, 40))
    act = np.ascontiguousarray((idd != 0).sum(0).astype(np.int32))
    P = _lib.PREC
    assert L.pbh_adjoint_choose_split(gs, ls, P["bf16x3"], None, 256) == -1
    assert L.pbh_adjoint_choose_split(gs, ls, P["fp16x3_peak"], act.ctypes.data, 256) == -5
    assert L.pbh_adjoint_choose_split(gs, ls, P["bf16x3"], act.ctypes.data, 256) == 7
    # the FP16 split kernel has the BF16 one's three workgroups per CU
    assert L.pbh_adjoint_choose_split(gs, ls, P["fp16x3"], act.ctypes.data, 256) == 7
    assert 1 <= L.pbh_adjoint_choose_split(gs, ls, P["fp16"], act.ctypes.data, 256) <= 16
    assert L.pbh_adjoint_choose_split(gs, ls, P["bf16"], act.ctypes.data, 256) == 10
    assert L.pbh_adjoint_choose_split(gs, ls, P["fp32"], act.ctypes.data, 256) == 5
    # every slice with all layers active on 8 CUs: more workgroups per slice only add
    # fixed work once the slots are full
    full = np.full(128, 40, np.int32)
    assert 1 <= L.pbh_adjoint_choose_split(gs, ls, P["fp32"], full.ctypes.data, 8) <= 2


def test_objective_adjoint_rejects_bad_arguments_without_launching():
    L = _lib.lib()
    g = _lib.PbhGrid(128, 128, 128, 0, 0.2, 0.2, 0.2, -12.7, -12.7, 0.0)
    lat = _lib.PbhLattice(41, 41, 40, 0, 0.5, 0.5, 0.0, 0.0)
    nul, one = C.c_void_p(0), C.c_void_p(256)
    P = _lib.PREC
    nb = L.pbh_objective_adjoint_workspace_bytes(g, lat, P["bf16x3"])
    # the adjoint's partials, R and the partial sums of J
    assert nb >= L.pbh_adjoint_workspace_bytes(g, lat, P["bf16x3"]) + 128 ** 3 * 4 + 1024 * 8
    args = lambda jout, prec, split, nbytes: (g, lat, one, one, one, one, one, nul, one, jout, prec, split,
                                              one, nbytes, nul)
    assert L.pbh_objective_adjoint(*args(nul, P["bf16x3"], 0, nb)) == -1          # no J output
    assert L.pbh_objective_adjoint(*args(one, P["fp16x3_peak"], 0, nb)) == -5     # no near-peak adjoint
    assert L.pbh_objective_adjoint(*args(one, P["bf16x3"], -1, nb)) == -2         # negative split
    assert L.pbh_objective_adjoint(*args(one, P["bf16x3"], 0, nb - 1)) == -7      # workspace too small
    assert L.pbh_objective_adjoint(*args(one, P["bf16x3"], 17, nb)) == -5         # split above 16
