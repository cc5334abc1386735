"""CPU oracle for the hedgehog_lpbf FD heat step (test infrastructure only).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.
"""
