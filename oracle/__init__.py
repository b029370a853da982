"""CPU oracle — TEST INFRASTRUCTURE ONLY.

A plain numpy/scipy restatement of the reference algorithms on frostspin's
block-sparse hot path: abelian SymmetricTensor storage, fusion, permute,
matmul, truncated SVD/eigh and find_block_cuts (abelian.py), SU(2) and O(2)
block layouts and permutes (su2.py, o2.py), the classical Ising C4v driver
(ising.py), and a numpy restatement of every device entry point of the
library for running the product's host logic on a CPU (device_emulator.py).
The CTMRG projector cut rule is restated in tests/test_projector_rule.py on top
of find_block_cuts. Every function cites the reference file:line it follows
(paths relative to the frostspin source tree).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker / CPU baseline. The product
(frostspin_amd) never imports it.

Pinning (which independent check fixes each function) is listed in each
module's docstring; see tests/test_oracle.py.
"""
