"""CPU oracle for the per-layer-exact Fermi-Eyges pencil-beam dose path.

TEST INFRASTRUCTURE ONLY. Nothing in the shipped ``pbdose`` package imports,
calls or links this directory: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or,
for ``bench.py``, as the timed CPU baseline), never as the thing measured or
shipped.

Modules
-------
pb_oracle      FP64 NumPy restatement of the reference physics LUT, problem
               model, separable engine, brute-force triple loop, error metric,
               gamma index and the adjoint of the weighted-quadratic objective.
torch_cpu      restatement of the reference's PyTorch batched-matmul engine,
               run on the host CPU (the reported CPU baseline).
"""
