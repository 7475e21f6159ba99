"""CPU oracle of the phase-field step: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this
package, and only as the checker (or the timed CPU baseline).  The product path never
calls it.  pf_oracle.py is the NumPy FP64 restatement of the reference; pf_oracle.c is
the same algorithm in C/OpenMP (FP64 and FP32) built into _build/libpforacle.so.
"""
