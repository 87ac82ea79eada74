"""Oracle for the ETD time-stepping path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, the smoke check of the entry module and the baseline legs of ``bench.py`` may
import this package, and only to check results against (or as the timed reference baseline).  The product package
``torchfsm_amd`` never imports it: it fails loudly when its HIP library is missing.

Modules:
  refpath  PyTorch restatement of the reference's schedule (full complex spectrum, the reference's
           op order and dtypes).  Pinned bit-for-bit on CPU against the golden fixtures written
           from the reference (tests/golden), and run on the GPU as the same-inputs oracle.
  numref   Independent NumPy float64 restatement (half spectrum, integer wavenumbers, phi
           functions from closed forms and Taylor series rather than contour means).  Pins
           refpath in float64 and is itself pinned by closed forms, brute-force DFTs, invariants
           and convergence orders.
"""
