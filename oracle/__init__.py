"""CPU restatement of xavs2's ME cost path (the checker).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product library never uses it.
"""
