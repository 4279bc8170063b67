"""ORACLE -- test infrastructure only.

A CPU restatement (numpy, float64/complex128) of the reference's BUG hot path, used as the checker
of the gfx950 library.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline leg may import it; the product package (``bondupdategalerkin_amd``) never does, and it
fails loudly when its HIP library is missing instead of falling back to anything here.

Every function cites the reference file:line it follows (paths relative to the reference
checkout, madhavmen/BondUpdateGalerkin).  How each part is pinned independently of itself (dense
contractions, exact diagonalisation, free fermions, closed forms, the reference's frozen
fixtures) is stated in the docstring of the module that holds it.
"""
