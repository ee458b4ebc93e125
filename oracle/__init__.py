"""FP64 host oracle for MEMD -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The HIP engine
(``memd-gpu_amd/``) never imports or links it.
"""
