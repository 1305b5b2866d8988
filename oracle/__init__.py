"""CPU oracle for the coupled-channel LDMD hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy restatement of the reference algorithm for the
exact coupled-channel path (W(R) contraction, Manolopoulos LDMD sweep, K->S
matching).  It exists to check the HIP product path; it is never part of it.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or as the labelled CPU
baseline).  The product package ``pyticc_amd`` must not import anything from
here and has no CPU fallback.

Every function cites the reference file:line it restates (paths relative to the
reference repository root).  How each function is pinned independently of its
own formula (closed forms, invariants, reference known-answer tests) is listed
in the module docstrings; a function with no such check says
"parity unpinned".
"""

from oracle import bessel, contraction, grid, ldmd, matching, pipeline  # noqa: F401
