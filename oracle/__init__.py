"""Oracle for the ERT forward + Jacobian path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import anything from here, and only as the checker or the CPU baseline, never
as the product path.

* ``oracle.ert_oracle`` -- this project's FP64 CPU restatement (numpy/scipy)
  of the reference algorithm, each function citing the reference file:line.
* ``oracle.ref`` -- ctypes access to the reference's own C++ ERT core,
  compiled in place from the pyGIMLi sources into ``oracle/_ref/``.
* ``oracle/oracle_port.c`` -- C restatement of the calc2 sensitivity loop with
  a thread pool, the CPU baseline for the Jacobian.
"""
