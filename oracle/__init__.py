"""CPU oracle for the 2-D TMz FDTD forward/adjoint path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker / CPU baseline.

* ``fdtd_oracle``   -- numpy restatement (FP64, plus an FP32 mode that follows
                       the reference's FP32 operation order).
* ``fdtd_oracle.c`` -- plain-C FP64 restatement of the same update equations
                       (OpenMP over shots), built by ``oracle/Makefile`` into
                       ``oracle/_build/liboracle.so``; used for the timed CPU
                       baseline and cross-checked against the numpy oracle.
"""
