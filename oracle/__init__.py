"""CPU oracle of the DiffPhD projective-dynamics path (TEST ORACLE ONLY).

Test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it, and only as the checker.
"""
