"""CPU oracle (NumPy, float64) for the matrix-product core.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package ``syngular_amd``
imports, calls or links anything in this directory.  The only permitted users
are ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``, and they use it purely as the checker (or, in bench.py, as the
timed CPU baseline), never as the thing that is measured or shipped.

Contents
--------
reference.py  A restatement, in code written for this repository, of the
              reference package's MatrixProductState / MatrixProductOperator
              semantics (tensor/matrix_product_state.py,
              tensor/matrix_product_operator.py, tensor/utils.py).  Householder
              QR comes from ``numpy.linalg.qr`` exactly as the reference uses it.
rounding.py   The textbook TT-rounding the reference does not have (left QR
              sweep, right SVD sweep with chi_max / relative cutoff, retained
              singular values and discarded weight per bond), the streamed
              transfer-matrix overlap, the materialised MPO x MPS site
              contraction and a few dense helpers used by the checks.

Pinning
-------
Every function is pinned by a check in ``tests/test_oracle.py`` that does not
retype its formula:
  * the known-answer pairs the reference's own test scripts print
    (test/test_mps.py, test/test_mpo.py, test/test_syn.py), turned into asserts;
  * dense closed forms: the QR-truncation projector ``A A^+`` applied to the
    dense tensor bond by bond, dense matrix products for ``@``, dense sums for
    ``+``, dense dot products for ``|``;
  * for the rounding: Eckart-Young at N=2 against ``numpy.linalg.svd`` of the
    dense matrix, Schmidt values of the dense tensor at the first truncated
    bond, the exact Pythagoras identity ||psi||^2 - ||psi'||^2 = sum_k w_k and
    brute force on tiny inputs.
No function here is "parity unpinned".
"""
