"""Test oracle for the TDSE + Hankel hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg, always as the checker and never as the thing measured or
shipped.  The product package (mma_amd) must not import it.
"""
