"""CPU oracle for the ADAMIXTURE likelihood hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
reference; the product path in ``adamixture_amd`` never imports it.
"""
from .oracle import *  # noqa: F401,F403
