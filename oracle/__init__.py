"""Test infrastructure only: CPU restatements of the reference ERNN path.

Nothing in the product (the ``ernn`` package, its HIP library or the
``bench.py`` timed GPU legs) imports, links or executes this package.  It is
used by ``tests/`` as the checker, by ``__graft_entry__.smoke()`` as the
checker, and by ``bench.py``'s ``cpu_baseline`` leg as the timed CPU port.

Modules
-------
ernn_fp64     FP64 NumPy restatement of the reference forward recurrence
              (Parameter_inversion/Inversion_adam.py:28-106, 240-307 and
              Damage_inversion/Inversion_adam.py:100-176, 331-393) plus the
              hand-derived discrete adjoint that replaces autograd
              (Parameter_inversion/Inversion_adam.py:447).
plain_torch   Op-for-op PyTorch restatement of the same reference code
              (F.pad + F.conv2d on transposed views, autograd through the
              whole time loop).  It is the FP32 trace/gradient parity target
              and the CPU baseline.

Parity pinning (what each oracle function is checked against) is listed in
the header of each module; see tests/test_oracle_*.py for the checks.
"""
