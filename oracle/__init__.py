"""CPU restatement of IQ-TREE 2's likelihood path (parity oracle, CPU baseline).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package as the checker; the product never does.
"""
