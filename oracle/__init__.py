"""CPU oracle for the particle-to-datacube path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``rubix_amd`` imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / the timed CPU baseline.

* ``ref_numpy``  - numpy restatement of the reference algorithm, op for op,
  with a selectable float type (float32 mirrors the reference's arithmetic,
  float64 is the high-precision oracle the errors are reported against).
* ``ref_torch``  - the same op graph in torch (CPU), used for autograd
  gradients (the stand-in for ``jax.grad`` of the reference) and as the
  timed CPU baseline of the plain per-step path.
"""
