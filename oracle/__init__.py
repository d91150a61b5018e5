"""oracle -- TEST INFRASTRUCTURE ONLY.

CPU checkers for the HIP BN254 G1 MSM:
  * oracle.bn254      pure-Python big-integer restatement (field, curve, GLV,
                      digits, closed form, seeded generators);
  * oracle.ref        ctypes binding of oracle/liboracle.so, the C restatement
                      of the reference CPU Pippenger (Sources/NeonFieldOps/bn254_msm.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product library (zkmetal_amd) never uses it.
"""
