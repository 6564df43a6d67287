"""Canonical algorithmic work of the path tracker (implementation
independent), used for the roofline figures.

Weights: complex multiply 6, complex add 2, complex FMA 8, complex divide
11; comparisons, moduli and square roots are not counted.

E  = one evaluation of H, H_z, H_t (src/homotopy.f90:188-231 over
     src/sparse.f90:223-280 with factors of one skipped):
       6 sum_j max(P_j - 1, 0)                    shared power table
     + sum_terms [ 6 max(nnz-1, 0) + 8            value c * prod
                 + 18 max(nnz-2, 0)               prefix/suffix products
                 + sum_{j: e_j>0} (8 + 6 [e_j>=2]) ]   derivatives
     + 6 n N + 94 n + 8 N                         homotopy assembly
LU(N)  = 8 sum_k (N-k)^2 + 11 sum_k (N-k)          src/linalg.f90:26-78
TRS(N) = 8 N (N-1) + 11 N                          src/linalg.f90:80-114

Per path: F = n_lin (E + LU + TRS) + n_be E, with n_lin the
linearizations (tangents + Newton steps) and n_be the backward-error
evaluations, both counted by the device per path.
"""
import numpy as np


def evaluation_flops(packed):
    n, N = packed.n, packed.n + 1
    h = packed.hexps
    P = h.max(axis=0) if len(h) else np.zeros(N, int)
    e = 6 * int(np.maximum(P - 1, 0).sum())
    nnz = (h > 0).sum(axis=1)
    e += int((6 * np.maximum(nnz - 1, 0) + 8 + 18 * np.maximum(nnz - 2, 0)).sum())
    e += int((8 * (h > 0) + 6 * (h >= 2)).sum())
    e += 6 * n * N + 94 * n + 8 * N
    return e


def lu_flops(N):
    k = np.arange(1, N + 1)
    return int(8 * ((N - k) ** 2).sum() + 11 * (N - k).sum())


def trs_flops(N):
    return 8 * N * (N - 1) + 11 * N


def unit_flops(packed):
    """(E, LU, TRS) of one linearization."""
    N = packed.n + 1
    return evaluation_flops(packed), lu_flops(N), trs_flops(N)


def path_flops(packed, nlin, nbe):
    E, LU, TRS = unit_flops(packed)
    return np.asarray(nlin, np.int64) * (E + LU + TRS) + np.asarray(nbe, np.int64) * E


# Bytes a launch must move at least: system program and constants, start
# points in, endpoint records out.
def algorithmic_bytes(packed, npaths):
    N = packed.n + 1
    program = packed.nterms * (16 + N * 4) + (N + 1) * 16 * 3
    per_path = 16 * N + (16 * N + 7 * 4 + 3 * 8)
    return program + npaths * per_path
