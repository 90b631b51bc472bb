"""Test oracles for the reconstruction path.

Test infrastructure only: the test suite, the smoke check and the CPU
baseline leg of the benchmark use these modules as the checker (or, for the
baseline, as the CPU implementation timed beside the GPU). The product
(pik_amd) never imports them.

  oracle.ref   ctypes binding of the reference's own decode-side code,
               compiled from its sources by oracle/Makefile into oracle/_ref
  oracle.port  ctypes binding of pik_oracle.c, the plain C restatement
"""
