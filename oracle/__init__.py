"""ORACLE — test infrastructure only.

CPU restatements of the reference algorithms (liboracle.so, built from the C++
sources in this directory).  Only tests/, the smoke check and the bench's
cpu_baseline leg may use this package; the product never does.
"""
