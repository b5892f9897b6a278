"""Test infrastructure: CPU oracle for the cluster-expansion Metropolis path.

Imported only by tests, the smoke check and the CPU-baseline leg of the
benchmark, and only as the checker.  The product never imports this package.
"""
