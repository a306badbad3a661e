"""CPU oracle for accusim's non-bonded force path (NumPy, float64).

TEST INFRASTRUCTURE ONLY. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import anything from this package, and
only as the checker or as the timed CPU baseline. The product package
(`devinproj_amd`) never imports it and has no CPU fallback.

Modules:
  accusim_np  restatement of the reference algorithm, function by function,
              each citing the reference file:line it follows.
  builders    seeded input builders for the workloads (FCC argon, rocksalt,
              Plummer sphere, uneven LJ batch).

Parity status of each function (how it is pinned) is listed in the header of
`accusim_np`.
"""
