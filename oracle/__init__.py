"""CPU oracle for the Vecchia / NNGP hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (`nngpr_amd`) imports,
links or executes anything under this directory.  It is used by `tests/`, by
`__graft_entry__.smoke()` and by the `cpu_baseline` leg of `bench.py`, always as
the checker or as the timed CPU baseline, never as the thing measured on the GPU.

Contents
--------
knn_oracle.c          FP64 brute-force kNN with the reference's key and tie rules
                      (built into liboracle_knn.so by build_oracle.py).
vecchia_oracle.py     FP64 restatement of fill -> Cholesky -> mean/std/log-likelihood
                      and of the reference's gradient formula.
numpy_port.py         Restatement of the reference's CPU path (generic kNN + per-matrix
                      LAPACK factor/solve over worker processes), used as the timed CPU
                      baseline.

Every function cites the reference file:line it follows.  How each function is
pinned (sklearn exact GP at m >= n-1, dense conditionals by brute force, central
finite differences, closed forms) is listed in tests/test_oracle.py.
"""
