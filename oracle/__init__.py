"""ORACLE -- test infrastructure only.

CPU restatement of the reference's hot path (Draco112358/Solver,
src/lib/solverFFT/).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only as the checker; the product
path (solver_amd) never imports it.
"""
