"""ctypes binding of libpbdose_hip.so (the C ABI in include/pbdose_hip.h).

The library is required: importing a function that needs it on a machine
without it raises. There is no CPU fallback anywhere in ``pbdose``.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpbdose_hip.so")

PBH_OK = 0
PBH_NOALLOC_SCRATCH_BYTES = 32 << 20      # include/pbdose_hip.h
PREC = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x3": 3, "fp16x3": 4,
        "bf16x3_peak": 5, "fp16x3_peak": 6}


class PbhGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("_pad", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double)]


class PbhLattice(C.Structure):
    _fields_ = [("n_sx", C.c_int32), ("n_sy", C.c_int32), ("n_layers", C.c_int32),
                ("_pad", C.c_int32), ("dx_spot", C.c_double), ("dy_spot", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double)]


class PbhBeamModel(C.Structure):
    _fields_ = [("sigma_x0", C.c_double), ("sigma_theta0", C.c_double), ("rho0", C.c_double),
                ("energy_spread", C.c_double), ("attenuation_length", C.c_double),
                ("fe_max_step", C.c_double), ("n_fine_idd", C.c_int32),
                ("fe_n_sub", C.c_int32), ("straggling", C.c_int32), ("straggling_depth", C.c_int32),
                ("icru_coef", C.c_double), ("icru_exp", C.c_double)]


class PbhBatchItem(C.Structure):
    _fields_ = [("grid", PbhGrid), ("lattice", PbhLattice),
                ("w", C.c_void_p), ("idd", C.c_void_p), ("sigma", C.c_void_p), ("dose", C.c_void_p),
                ("presc", C.c_void_p), ("vox_w", C.c_void_p), ("j_out", C.c_void_p),
                ("grad_w", C.c_void_p), ("grad_idd", C.c_void_p), ("grad_sigma", C.c_void_p)]


class PbhBatchTuning(C.Structure):
    _fields_ = [("fwd_chunk", C.c_int32), ("adj_split", C.c_int32), ("group", C.c_int32),
                ("order", C.c_int32), ("tile_rows", C.c_int32), ("persistent", C.c_int32)]


_P = C.c_void_p
_G = C.POINTER(PbhGrid)
_L = C.POINTER(PbhLattice)
_M = C.POINTER(PbhBeamModel)

# name -> (restype, argtypes); every symbol include/pbdose_hip.h declares.
SIGNATURES = {
    "pbh_strerror": (C.c_char_p, [C.c_int]),
    "pbh_version": (C.c_char_p, []),
    "pbh_validate": (C.c_int, [_G, _L]),
    "pbh_range_table_size": (C.c_int, []),
    "pbh_range_table": (C.c_int, [_P, _P, _P]),
    "pbh_interp_table": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, _P]),
    "pbh_interp_table_slope": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, C.c_int, _P]),
    "pbh_lut_workspace_bytes": (C.c_size_t, [_M, C.c_int, C.c_double, C.c_double, C.c_int]),
    "pbh_build_lut": (C.c_int, [_M, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P, _P, _P, _P, _P,
                                _P, C.c_size_t, _P]),
    "pbh_lut_tangent_workspace_bytes": (C.c_size_t, [_M, C.c_int, C.c_double, C.c_double, C.c_int]),
    "pbh_build_lut_tangent": (C.c_int, [_M, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P, _P, _P,
                                        _P, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "pbh_forward_workspace_bytes": (C.c_size_t, [_G, _L, C.c_int]),
    "pbh_forward_workspace_bytes_tuned": (C.c_size_t, [_G, _L, C.c_int, C.c_int]),
    "pbh_forward": (C.c_int, [_G, _L, _P, _P, _P, _P, C.c_int, _P, C.c_size_t, _P]),
    "pbh_forward_tuned": (C.c_int, [_G, _L, _P, _P, _P, _P, C.c_int, C.c_int, _P, C.c_size_t, _P]),
    "pbh_forward_choose_chunk": (C.c_int, [_G, _L, C.c_int, _P, C.c_int]),
    "pbh_forward_profiled": (C.c_int, [_G, _L, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, C.c_size_t,
                                       _P, _P, _P]),
    "pbh_forward_choose_shape": (C.c_int, [_G, _L, C.c_int, _P, C.c_int, _P]),
    "pbh_forward_shaped": (C.c_int, [_G, _L, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, C.c_size_t, _P]),
    "pbh_forward_workspace_bytes_shaped": (C.c_size_t, [_G, _L, C.c_int, C.c_int, C.c_int]),
    "pbh_forward_objective_adjoint_workspace_bytes": (C.c_size_t, [_G, _L, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pbh_forward_objective_adjoint": (C.c_int, [_G, _L, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                                C.c_int, C.c_int, _P, C.c_size_t, _P]),
    "pbh_forward_fp64": (C.c_int, [_G, _L, _P, _P, _P, _P, _P]),
    "pbh_average_sigma": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "pbh_average_sigma_fp64": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "pbh_forward_noalloc": (C.c_int, [_G, _L, _P, _P, _P, _P, C.c_int, C.c_int, _P]),
    "pbh_adjoint_workspace_bytes": (C.c_size_t, [_G, _L, C.c_int]),
    "pbh_adjoint": (C.c_int, [_G, _L, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P, C.c_size_t, _P]),
    "pbh_adjoint_tuned": (C.c_int, [_G, _L, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_int, _P, C.c_size_t, _P]),
    "pbh_adjoint_choose_split": (C.c_int, [_G, _L, C.c_int, _P, C.c_int]),
    "pbh_objective_workspace_bytes": (C.c_size_t, [_G]),
    "pbh_objective": (C.c_int, [_G, _P, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "pbh_objective_adjoint_workspace_bytes": (C.c_size_t, [_G, _L, C.c_int]),
    "pbh_objective_adjoint": (C.c_int, [_G, _L, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_int, _P,
                                        C.c_size_t, _P]),
    "pbh_lut_chain": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, _P, _P]),
    "pbh_batch_workspace_bytes": (C.c_size_t, [C.POINTER(PbhBatchItem), C.c_int, C.c_int,
                                               C.POINTER(PbhBatchTuning)]),
    "pbh_batch_create": (C.c_int, [C.POINTER(PbhBatchItem), C.c_int, C.c_int, C.POINTER(PbhBatchTuning),
                                   _P, C.c_size_t, _P, C.POINTER(C.c_void_p)]),
    "pbh_batch_forward_adjoint": (C.c_int, [_P, _P]),
    "pbh_batch_launch_count": (C.c_int, [_P]),
    "pbh_batch_destroy": (None, [_P]),
    "pbh_weight_step": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_size_t, C.c_int, C.c_float, C.c_float,
                                  C.c_float, _P]),
    "pbh_gamma_workspace_bytes": (C.c_size_t, [_G]),
    "pbh_gamma_count_offsets": (C.c_int, [_G, C.c_double]),
    "pbh_gamma_index": (C.c_int, [_G, _P, C.c_int, _P, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int, _P, _P, _P, _P, C.c_size_t, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library (loaded once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(pbdose has no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _lib = h
    return _lib


def strerror(rc: int) -> str:
    return lib().pbh_strerror(rc).decode()


def check(rc: int, what: str) -> None:
    if rc != PBH_OK:
        raise RuntimeError(f"{what} failed: {strerror(rc)} (code {rc})")


def require_cuda(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError(f"pbdose runs on the GPU only (got device={device!r}); "
                         "there is no CPU path")
    if not torch.cuda.is_available():
        raise RuntimeError("pbdose needs a ROCm GPU, and torch.cuda.is_available() is False")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def stream_of(dev: torch.device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def ptr(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def workspace(nbytes: int, dev: torch.device) -> torch.Tensor:
    """Caller-owned, zero-filled scratch for the C ABI (the library never allocates)."""
    return torch.zeros(max(int(nbytes), 1), dtype=torch.uint8, device=dev)


def grid_struct(g) -> PbhGrid:
    return PbhGrid(int(g.nx), int(g.ny), int(g.nz), 0, float(g.dx), float(g.dy), float(g.dz),
                   float(g.ox), float(g.oy), float(g.oz))


def lattice_struct(lat) -> PbhLattice:
    return PbhLattice(int(lat.n_sx), int(lat.n_sy), int(lat.n_layers), 0,
                      float(lat.dx_spot), float(lat.dy_spot), float(lat.x0), float(lat.y0))
