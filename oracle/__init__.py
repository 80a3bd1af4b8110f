"""CPU FP64 oracle of the radiation path (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU baseline, as the checker.
"""
