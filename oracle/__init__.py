"""The oracle: CPU restatement of dim's CSR5 path.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product package.
"""
