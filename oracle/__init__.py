"""ORACLE -- test infrastructure only.

A CPU restatement of the Instant-angelo NeuS / Neuralangelo hot path
(hugoycj/Instant-angelo, models/*.py, with the tiny-cuda-nn and nerfacc 0.3.3
semantics those files call).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import, call or link anything here, and only
as the checker / the CPU baseline.  The product (instant-angelo_amd/) never
imports it and has no CPU fallback.

Layout:
  oracle.c        scalar C, bit-exact fp32 restatement of the integer / index /
                  position arithmetic (level table, hash-grid corners and
                  encode, ray-AABB, occupancy marching).
  clib.py         builds / loads oracle.c (ctypes).
  hashgrid.py     level table, exact encode (C), differentiable encode (torch).
  nets.py         VanillaMLP (weight-norm, sphere init, Softplus/ReLU), SH.
  field.py        VolumeSDF.forward incl. finite differences and curvature.
  render.py       NeuS alpha, transmittance, accumulation, output dict, losses.
  march.py        ray-AABB + occupancy-grid marching (C) and occupancy update.
  scene.py        seeded cameras, rays, analytic scenes, uniform samples.

Pinning (what fixes each function independently of its own formula) is
listed per function and exercised in tests/test_oracle_*.py.  Functions with
no independent pin say "parity unpinned" in their docstring.
"""
