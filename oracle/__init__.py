"""CPU oracle — TEST INFRASTRUCTURE ONLY.

A plain numpy/scipy restatement of the reference algorithms on frostspin's
block-sparse hot path (abelian SymmetricTensor storage, fusion, permute,
matmul, truncated SVD/eigh, find_block_cuts, CTMRG projectors and moves, the
classical Ising C4v driver). Every function cites the reference file:line it
follows (paths relative to the frostspin source tree).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker / CPU baseline. The product
(frostspin_amd) never imports it.

Pinning (which independent check fixes each function) is listed in each
module's docstring; see tests/test_oracle.py.
"""
