"""Reader and writer for the reference's ``inout.txt`` data files.

Reading follows ReadIOFile (src/workproc.f90:53-424): the keyword header,
the history line of the last function, the Basis Building and Optimisation
Program (BBOP) steps, the history of every function and one line
``j  vechL_j  vechB_j`` per basis function; Fortran ``D`` exponents are
accepted.  Writing follows SaveResults (src/workproc.f90:428-569) field for
field and format for format (``D23.16`` reals, ``I6``/``I7`` integers), so a
file read and written back is identical to the one the reference wrote.
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import List, Optional

import numpy as np

from .system import EcgSystem, make_system

__all__ = ["InoutFile", "BbopStep", "HistoryEntry", "read_inout", "write_inout", "format_inout", "fortran_d"]

SEPARATOR = " " + "=" * 30

# per BBOP action: its fields in file order (i: integer, r: real, f: file
# name, m: the GSEP method letter), as ReadIOFile reads and SaveResults
# writes them
_BBOP_FIELDS = {
    "BASIS_ENL": "m iiiii rr",
    "OPT_CYCLE": "m iiiiiii rr",
    "FULL_OPT1": "m iiii rr ii f",
    "EXPC_VALS": "m i",
    "DENSITIES": "m i ffff",
    "ELIM_LCFN": "m i r f",
    "ELIM_LND1": "m i r f",
    "SEPR_LND1": "m i rr f",
    "SEPR_FLCF": "m i rr f",
    "SAVE_FILE": "i f",
}

_TOKEN = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+)(?:[DdEe][-+]?\d+)?|[-+]?\d+|\S+")


def _f(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def fortran_d(x: float, w: int = 23, d: int = 16) -> str:
    """Fortran ``Dw.d`` edit descriptor: 0.dddd...D+ee, right-justified in w
    columns (the exponent letter is dropped for three-digit exponents)."""
    if x == 0.0 or not math.isfinite(x):
        body = "0." + "0" * d + "D+00" if x == 0.0 else str(x)
        return body.rjust(w)
    m, e = ("%.*E" % (d - 1, abs(x))).split("E")
    e = int(e) + 1
    digits = m.replace(".", "")
    if abs(e) <= 99:
        exp = "D%s%02d" % ("+" if e >= 0 else "-", abs(e))
    else:
        exp = "%s%03d" % ("+" if e >= 0 else "-", abs(e))
    body = ("-" if x < 0 else "") + "0." + digits + exp
    return body.rjust(w)


@dataclasses.dataclass
class BbopStep:
    action: str
    method: str = ""
    ints: List[int] = dataclasses.field(default_factory=list)     # A, B, C, ...
    reals: List[float] = dataclasses.field(default_factory=list)  # Q, R
    files: List[str] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class HistoryEntry:
    energy: float
    cycles_done: int = 0
    init_func_at_last_step: int = 0
    num_energy_evals_full_opt: int = 0


@dataclasses.dataclass
class InoutFile:
    system: EcgSystem
    basis_size: int
    current_energy: float
    which_eigenvalue: int
    eigval_tolerance: float
    invit_parameter: float
    params: np.ndarray          # (K, npt) float64, row j = (vech L_j, vech B_j)
    history_energy: np.ndarray  # (K,) energy after function j was added
    masses: Optional[List[float]] = None
    charges: Optional[List[float]] = None
    last_eigval_tol: float = 0.0
    best_eigval_tol: float = 0.0
    worst_eigval_tol: float = 0.0
    generator_param: List[float] = dataclasses.field(default_factory=lambda: [0.0, 0.0, 0.0])
    bbop: List[BbopStep] = dataclasses.field(default_factory=list)
    history: List[HistoryEntry] = dataclasses.field(default_factory=list)


def _parse_bbop(line: str) -> BbopStep:
    toks = _TOKEN.findall(line)
    action = toks[0]
    spec = _BBOP_FIELDS[action].replace(" ", "")
    st = BbopStep(action=action)
    k = 1
    for c in spec:
        t = toks[k]
        k += 1
        if c == "m":
            st.method = t
        elif c == "i":
            st.ints.append(int(t))
        elif c == "r":
            st.reals.append(_f(t))
        else:
            st.files.append(t)
    return st


def _format_bbop(st: BbopStep) -> str:
    a = st.action
    ints, reals, files = st.ints, st.reals, st.files
    i6 = lambda v: "%6d" % v
    if a in ("BASIS_ENL", "OPT_CYCLE"):
        return " %s %s%s %s %s" % (a, st.method, "".join(" " + i6(v) for v in ints), fortran_d(reals[0]),
                                   fortran_d(reals[1]))
    if a == "FULL_OPT1":
        return " %s %s%s %s %s %s %s %s" % (a, st.method, "".join(" " + i6(v) for v in ints[:4]),
                                            fortran_d(reals[0]), fortran_d(reals[1]), i6(ints[4]), i6(ints[5]),
                                            files[0])
    if a == "EXPC_VALS":
        return " %s %s %s" % (a, st.method, i6(ints[0]))
    if a == "DENSITIES":
        return " %s %s %s %s" % (a, st.method, i6(ints[0]), " ".join(files[:4]))
    if a in ("ELIM_LCFN", "ELIM_LND1"):
        return " %s %s %s%s %s" % (a, st.method, i6(ints[0]), fortran_d(reals[0]), files[0])
    if a in ("SEPR_LND1", "SEPR_FLCF"):
        return " %s %s %s%s %s %s" % (a, st.method, i6(ints[0]), fortran_d(reals[0]), fortran_d(reals[1]),
                                      files[0])
    if a == "SAVE_FILE":
        return " %s %s %s" % (a, i6(ints[0]), files[0])
    raise ValueError("unknown BBOP action %r" % a)


def _history_line(j: int, h: HistoryEntry) -> str:
    return " %6d %s %6d %6d %7d" % (j, fortran_d(h.energy), h.cycles_done, h.init_func_at_last_step,
                                    h.num_energy_evals_full_opt)


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
    bbop = [_parse_bbop(lines[j]) for j in range(seps[1] + 1, seps[2]) if lines[j].strip()]
    history = []
    j = seps[2] + 1
    for _ in range(K):
        toks = lines[j].split()
        history.append(HistoryEntry(_f(toks[1]), int(toks[2]), int(toks[3]), int(toks[4])))
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
    gp = [_f(t) for t in head.get("GENERATOR_PARAM", ["0", "0", "0"])[:3]]
    return InoutFile(
        system=system,
        basis_size=K,
        current_energy=_f(head["CURRENT_ENERGY"][0]),
        which_eigenvalue=int(head.get("WHICH_EIGENVALUE", ["1"])[0]),
        eigval_tolerance=_f(head.get("EIGVAL_TOLERANCE", ["1D-12"])[0]),
        invit_parameter=_f(head.get("INVITPARAMETER", ["1D0"])[0]),
        params=P,
        history_energy=np.array([h.energy for h in history]),
        masses=masses,
        charges=charges,
        last_eigval_tol=_f(head.get("LAST_EIGVAL_TOL", ["0"])[0]),
        best_eigval_tol=_f(head.get("BEST_EIGVAL_TOL", ["0"])[0]),
        worst_eigval_tol=_f(head.get("WORST_EIGVAL_TOL", ["0"])[0]),
        generator_param=gp,
        bbop=bbop,
        history=history,
    )


def format_inout(f: InoutFile, func_num=None) -> str:
    """The text SaveResults writes for f.  Without func_num the parameter
    lines follow the order of f.params; with it (1-based, func_num[k] the
    line of function k, like Glob_FuncNum) they are written in that order,
    as SaveResults(Sort='yes') does (src/workproc.f90:552-563).  History
    lines keep their order either way."""
    K = f.params.shape[0]
    rows = np.arange(K)
    if func_num is not None:
        fn = np.asarray(func_num, dtype=np.int64) - 1
        if fn.shape != (K,) or not np.array_equal(np.sort(fn), rows):
            raise ValueError("func_num must be a permutation of 1..K")
        rows = np.empty(K, np.int64)
        rows[fn] = np.arange(K)
    masses = f.masses if f.masses is not None else list(f.system.masses)
    charges = f.charges if f.charges is not None else list(f.system.charges)
    hist = f.history if f.history else [HistoryEntry(float(e)) for e in f.history_energy]
    if len(hist) != K:
        raise ValueError("history has %d entries for %d functions" % (len(hist), K))
    out = [
        " PARTICLES%3d" % len(masses),
        " MASSES" + "".join(" " + fortran_d(m) for m in masses),
        " CHARGES" + "".join(" " + fortran_d(q) for q in charges),
        " SYMMETRY " + f.system.young.strip(),
        " BASIS_SIZE %6d" % K,
        " CURRENT_ENERGY   " + fortran_d(f.current_energy),
        " WHICH_EIGENVALUE %6d" % f.which_eigenvalue,
        " EIGVAL_TOLERANCE " + fortran_d(f.eigval_tolerance),
        " INVITPARAMETER   " + fortran_d(f.invit_parameter),
        " LAST_EIGVAL_TOL  " + fortran_d(f.last_eigval_tol),
        " BEST_EIGVAL_TOL  " + fortran_d(f.best_eigval_tol),
        " WORST_EIGVAL_TOL " + fortran_d(f.worst_eigval_tol),
        " GENERATOR_PARAM" + "".join(" " + fortran_d(g) for g in f.generator_param),
        SEPARATOR,
        _history_line(K, hist[K - 1]),
        SEPARATOR,
    ]
    out += [_format_bbop(st) for st in f.bbop]
    out.append(SEPARATOR)
    out += [_history_line(j + 1, hist[j]) for j in range(K)]
    out.append(SEPARATOR)
    out += [" %6d" % (j + 1) + "".join(" " + fortran_d(v) for v in f.params[rows[j]]) for j in range(K)]
    return "\n".join(out) + "\n"


def write_inout(path, f: InoutFile, func_num=None) -> None:
    with open(path, "w") as fh:
        fh.write(format_inout(f, func_num))
