"""CPU float64 oracle for the streaming LFMM2 hot path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import it.
The product path (``schnellfmm_amd``) never imports anything from here and
fails loudly when its HIP library is missing.
"""
from .lfmm_oracle import *  # noqa: F401,F403
