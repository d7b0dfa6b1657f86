"""Process-wide trace settings (batoid/global_vars.py:1-9).

``_batoid_niter`` is the fixed number of Newton iterations every
intersection runs; it is part of the trace's semantics.  The reference's
``_batoid_max_threads`` (OpenMP threads) has no role on the GPU and is kept
only so code that reads it still works."""
import os

_batoid_max_threads = int(os.environ.get('OMP_NUM_THREADS', 1))
_batoid_niter = 5
