"""Oracle: a CPU (NumPy) restatement of the reference IB-LBM algorithm.

This package is TEST INFRASTRUCTURE. It is the checker against which the HIP
implementation in ``vivsim_amd`` is compared. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path never routes through it.

Every function restates one function of the reference package
(haimingz/vivsim, pure JAX) and cites the file:line it follows. The reference
cannot run here (JAX is absent and its sources need Python >= 3.11), so the
oracle is pinned by closed forms, conservation invariants, brute-force loops
and the identities listed in ``tests/test_oracle_*.py``.

Precision: every function computes in the dtype of its floating inputs, so the
same code serves as the FP64 "truth" and as the FP32 mirror of the reference's
default (x32) JAX arithmetic.
"""
