"""Reader for the reference's ``inout.txt`` data files (the subset the pair
path needs: system, symmetry, basis size, current energy, target root,
inverse-iteration settings and the nonlinear parameters).

Format followed: src/workproc.f90:53-184 (keyword header) and :375-424
(history block, then one line ``j  vechL_j  vechB_j`` per basis function).
Fortran ``D`` exponents are accepted.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .system import EcgSystem, make_system

__all__ = ["InoutFile", "read_inout"]


def _f(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


@dataclasses.dataclass
class InoutFile:
    system: EcgSystem
    basis_size: int
    current_energy: float
    which_eigenvalue: int
    eigval_tolerance: float
    invit_parameter: float
    params: np.ndarray        # (K, npt) float64, row j = (vech L_j, vech B_j)
    history_energy: np.ndarray  # (K,) energy after function j was added


def read_inout(path) -> InoutFile:
    with open(path, "r") as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    head = {}
    i = 0
    while i < len(lines) and not lines[i].strip().startswith("===="):
        toks = lines[i].split()
        if toks:
            head[toks[0]] = toks[1:]
        i += 1
    for key in ("PARTICLES", "MASSES", "CHARGES", "SYMMETRY", "BASIS_SIZE", "CURRENT_ENERGY"):
        if key not in head:
            raise ValueError("%s: missing %s line" % (path, key))
    npart = int(head["PARTICLES"][0])
    masses = [_f(t) for t in head["MASSES"][:npart]]
    charges = [_f(t) for t in head["CHARGES"][:npart]]
    young = "".join(head["SYMMETRY"])
    K = int(head["BASIS_SIZE"][0])
    system = make_system(masses, charges, young)
    npt = system.npt
    seps = [j for j, ln in enumerate(lines) if ln.strip().startswith("====")]
    if len(seps) < 4:
        raise ValueError("%s: expected four separator lines" % path)
    hist = []
    j = seps[2] + 1
    for _ in range(K):
        toks = lines[j].split()
        hist.append(_f(toks[1]))
        j += 1
    params = []
    j = seps[3] + 1
    while len(params) < K:
        toks = lines[j].split()
        j += 1
        if not toks:
            continue
        params.append([_f(t) for t in toks[1:1 + npt]])
    P = np.array(params, dtype=np.float64)
    if P.shape != (K, npt):
        raise ValueError("%s: parameter block has shape %s" % (path, P.shape))
    return InoutFile(
        system=system,
        basis_size=K,
        current_energy=_f(head["CURRENT_ENERGY"][0]),
        which_eigenvalue=int(head.get("WHICH_EIGENVALUE", ["1"])[0]),
        eigval_tolerance=_f(head.get("EIGVAL_TOLERANCE", ["1D-12"])[0]),
        invit_parameter=_f(head.get("INVITPARAMETER", ["1D0"])[0]),
        params=P,
        history_energy=np.array(hist),
    )
