"""pbdose -- per-layer-exact Fermi-Eyges pencil-beam dose on MI355X (gfx950).

Same module names as the reference package (``physics``, ``model``,
``engines``, ``gamma``); the dose, look-up-table, adjoint, robust-batch and gamma
arithmetic runs in the hand-written HIP kernels of ``libpbdose_hip.so``.
``planning`` (objective, adjoint, autograd) and ``robust`` (scenario batches)
have no counterpart in the reference.
"""
__version__ = "0.1.0"
__all__ = ["physics", "model", "engines", "gamma", "planning", "robust"]
