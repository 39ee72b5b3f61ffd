"""CPU oracle of the Barnes-Hut step (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  Never part of the product path.
"""
