"""FP64 CPU oracle for the dense GP core -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy FP64, the operation sequence of the
reference R package gplite (v0.13.0) for the dense ``method_full`` path:
covariance builds, the B = I + S K S Cholesky chain, the Laplace Newton loop,
parallel EP, the log evidence, the predictive mean/variance, and the
hyperparameter gradient the reference lacks.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it, and only as the checker (or the timed CPU baseline).  The
product path (``gplite_amd``) never imports it.

Pinning: every function is checked in ``tests/test_oracle.py`` against closed
forms, scipy/LAPACK, numpy's Gauss-Hermite rule, brute-force loops or finite
differences (see each function's docstring for which).
"""

from .gp_oracle import *  # noqa: F401,F403
