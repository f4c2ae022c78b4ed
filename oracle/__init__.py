"""CPU oracle for the barotropic RK2 step — TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / CPU baseline, never as the thing measured.

* ``barotr_np``  — vectorised NumPy FP64 restatement of the reference's JAX
  step, one function per reference function (file:line cited on each).
* ``loop_ref``   — independent per-point loop restatement from the closed
  index forms, for tiny grids (C4–C8), used to cross-check ``barotr_np``.
"""
