"""Host-CPU restatement of the reference's PyTorch batched-matmul engine.

TEST INFRASTRUCTURE ONLY: tests use it as a checker, and bench.py times it as the
reported CPU comparison ("port"); the shipped package never imports it.

Same op sequence as engines.py:256-274 (per_layer=True): broadcast exp
factors Gx, Gy for every (layer, depth), A = IDD x w, T = bmm(A^T, Gy^T),
Dl = bmm(Gx, T), D = sum over layers, transposed to (nz, ny, nx).
"""
from __future__ import annotations

import numpy as np
import torch

INV_SQRT_2PI = 1.0 / np.sqrt(2.0 * np.pi)


def torch_separable_cpu(x, y, sx, sy, w, idd, sigma, dtype=torch.float64, layer_chunk=None):
    xv, yv = torch.as_tensor(x, dtype=dtype), torch.as_tensor(y, dtype=dtype)
    xs, ys = torch.as_tensor(sx, dtype=dtype), torch.as_tensor(sy, dtype=dtype)
    w = torch.as_tensor(w, dtype=dtype)
    idd = torch.as_tensor(idd, dtype=dtype)
    sig = torch.as_tensor(sigma, dtype=dtype)
    L, nz = idd.shape
    nx, ny, nsx, nsy = len(x), len(y), len(sx), len(sy)
    D = torch.zeros((nz, ny, nx), dtype=dtype)
    dx2 = (xv.view(1, 1, -1, 1) - xs.view(1, 1, 1, -1)) ** 2
    dy2 = (yv.view(1, 1, -1, 1) - ys.view(1, 1, 1, -1)) ** 2
    step = layer_chunk or L
    for l0 in range(0, L, step):
        l1 = min(L, l0 + step)
        n = l1 - l0
        s = sig[l0:l1].clamp_min(1e-9).reshape(n, nz, 1, 1)
        norm = INV_SQRT_2PI / s
        Gx = (norm * torch.exp(-dx2 / (2.0 * s * s))).reshape(n * nz, nx, nsx)
        Gy = (norm * torch.exp(-dy2 / (2.0 * s * s))).reshape(n * nz, ny, nsy)
        A = (idd[l0:l1, :, None, None] * w[l0:l1, None]).reshape(n * nz, nsy, nsx)
        T = torch.bmm(A.transpose(1, 2), Gy.transpose(1, 2))
        Dl = torch.bmm(Gx, T)
        D += Dl.reshape(n, nz, nx, ny).sum(0).transpose(1, 2)
    return D
