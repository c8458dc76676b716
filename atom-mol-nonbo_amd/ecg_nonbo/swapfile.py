"""The reference's binary swap file of H and S.

StoreMatricesInSwapFile (src/workproc.f90:2251-2311) writes, with
``form='binary'`` (a stream without record markers), a 4-byte integer K,
then K columns of one K x K complex128 matrix W, column-major, and then the
K doubles of diagS.  W holds H on and below the diagonal and S strictly
above it: W(j,i) = S(i,j) for j < i.  With the shift-and-invert layout
('I': the lower triangle holds H - sigma S and S has a unit diagonal) the
shift is taken out before writing, so a file never depends on sigma.
ReadSwapFileAndDistributeData (src/workproc.f90:2158-2248) accepts a file
only when its K equals the current basis size, restores the layout of the
method it is asked for, and afterwards overwrites the file with the text
``Swap file is empty`` so that stale matrices are not read twice.

Arrays are those of ComputeMatElem (Fortran-ordered complex (K, K) H and S
with the lower triangles filled, diagH, diagS), as both the HIP context's
``ComputeMatElem`` and the oracle return them.
"""
from __future__ import annotations

import os
from typing import Optional, Tuple

import numpy as np

__all__ = ["write_swap", "read_swap", "EMPTY_MARK"]

EMPTY_MARK = b"Swap file is empty"


def _check_method(method: str) -> str:
    m = method.upper()
    if m not in ("G", "I"):
        raise ValueError("method must be 'G' or 'I', not %r" % method)
    return m


def write_swap(path, H, S, diagH, diagS, method: str = "G", approx_energy: float = 0.0) -> None:
    """Write H and S of a K-function basis (K = len(diagS)) in the swap layout."""
    m = _check_method(method)
    diagS = np.asarray(diagS, dtype=np.float64)
    K = diagS.shape[0]
    H = np.asarray(H)[:K, :K]
    S = np.asarray(S)[:K, :K]
    low = np.tril_indices(K, -1)
    W = np.zeros((K, K), np.complex128)
    if m == "G":
        W[low] = H[low]
        W[np.diag_indices(K)] = np.asarray(diagH, dtype=np.float64)[:K]
    else:
        # H - sigma S (unit diagonal of S) back to H
        W[low] = H[low] + approx_energy * S[low]
        W[np.diag_indices(K)] = np.diagonal(H) + approx_energy
    W[low[1], low[0]] = S[low]
    with open(path, "wb") as fh:
        fh.write(np.int32(K).tobytes())
        fh.write(W.tobytes(order="F"))
        fh.write(diagS.tobytes())


def read_swap(path, K: int, method: str = "G", approx_energy: float = 0.0,
              erase: bool = True) -> Optional[Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]]:
    """(H, S, diagH, diagS) in the layout of ``method`` if ``path`` holds the
    matrices of a K-function basis, else None.  A file whose K matches is
    overwritten with the empty mark after reading, whether or not it was
    complete (``erase=False`` keeps it)."""
    m = _check_method(method)
    if not os.path.isfile(path):
        return None
    with open(path, "rb") as fh:
        head = fh.read(4)
        if len(head) < 4 or int(np.frombuffer(head, np.int32)[0]) != K:
            return None
        body = fh.read(16 * K * K + 8 * K)
    if erase:
        with open(path, "wb") as fh:
            fh.write(EMPTY_MARK)
    if len(body) != 16 * K * K + 8 * K:
        return None
    W = np.frombuffer(body, np.complex128, K * K).reshape((K, K), order="F")
    diagS = np.frombuffer(body, np.float64, K, offset=16 * K * K).copy()
    low = np.tril_indices(K, -1)
    H = np.zeros((K, K), np.complex128, order="F")
    S = np.zeros((K, K), np.complex128, order="F")
    S[low] = W[low[1], low[0]]
    diagH = np.ascontiguousarray(W.diagonal().real)
    if m == "G":
        H[low] = W[low]
        H[np.diag_indices(K)] = W.diagonal()
    else:
        S[low[1], low[0]] = np.conj(S[low])
        S[np.diag_indices(K)] = 1.0
        H[np.diag_indices(K)] = W.diagonal() - approx_energy
        H[low] = W[low] - approx_energy * S[low]
    return H, S, diagH, diagS
