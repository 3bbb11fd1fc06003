"""TEST INFRASTRUCTURE ONLY: the CPU oracle for the static fill and the loop/port solve.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker (never as the thing measured or shipped).

  oracle.fill   FillGeometry.prepare / kinds / far and analytic pair rules in numpy,
                plus ctypes access to libcpu_ref.so (cpu_ref.c, the C restatement of
                the reference's static_fill / static_pairs / static_combine) and to the
                reference's own compiled core built into oracle/_ref (when present).
  oracle.mqs    RWG basis, sheet resistance, via R/L, loop basis, dense port solve.
  oracle.gmres  the plain-preconditioned lockstep GMRES of the port solve (the timed CPU
                GMRES baseline).
  oracle.biot   Biot-Savart of RWG sheet currents and of straight filaments.
  oracle.mpnn   the message-passing network of the learned correction, FP64 with the
                device's BF16 rounding points.

Parity status of each function is stated in its docstring (what pins it).
"""
