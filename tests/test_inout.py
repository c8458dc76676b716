"""inout.txt reader/writer, BBOP parser and swap file (no GPU).

  * the six HD+ data files the reference ships, read and written back, are
    byte-identical to the files (every header field, BBOP line, history line
    and parameter passes through the writer's formats);
  * the D23.16 formatter against a Fortran compiler's own D23.16 output on
    edge cases and random doubles (skipped when no flang is installed);
  * the BBOP actions the fixtures do not use, against line images taken from
    SaveResults' format strings (src/workproc.f90:490-538);
  * the swap file's byte layout, its round trip in both storage modes and the
    reference's acceptance rules (K must match, erased after reading).
"""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

from ecg_nonbo import read_inout
from ecg_nonbo.inout import BbopStep, HistoryEntry, _format_bbop, _parse_bbop, format_inout, fortran_d, write_inout
from ecg_nonbo.swapfile import EMPTY_MARK, read_swap, write_swap
from oracle import oracle

FIXTURES = ["inout_1_500", "inout_2_500", "inout_1_1000", "inout_2_1000", "inout_1_1500", "inout_2_1500"]


def _fixture(golden_dir, name):
    return os.path.join(golden_dir, "hdplus", name + ".txt")


@pytest.mark.parametrize("fname", FIXTURES)
def test_fixture_round_trip_is_byte_identical(golden_dir, fname, tmp_path):
    path = _fixture(golden_dir, fname)
    with open(path, "rb") as fh:
        raw = fh.read()
    f = read_inout(path)
    out = tmp_path / "inout.txt"
    write_inout(str(out), f)
    assert out.read_bytes() == raw


def test_fixture_fields(golden_dir):
    f = read_inout(_fixture(golden_dir, "inout_1_1500"))
    assert f.masses == [3670.4829652, 1836.15267261, 1.0]
    assert f.charges == [1.0, 1.0, -1.0]
    assert f.system.young.strip() == "(P11)"
    assert (f.last_eigval_tol, f.best_eigval_tol, f.worst_eigval_tol) == (0.12, 0.11, 1.4e-20)
    assert f.generator_param == [0.7, 1.0, 3.0]
    assert len(f.bbop) == 13
    st = f.bbop[0]
    assert (st.action, st.method, st.ints, st.reals) == ("BASIS_ENL", "G", [1400, 1420, 1, 500, 200], [0.99, 50.0])
    st = f.bbop[-1]
    assert (st.action, st.ints) == ("OPT_CYCLE", [1500, 1, 1500, 1, 1, 70, 120])
    assert f.history[-1] == HistoryEntry(-0.5978979685879846, 57, 487, 0)
    assert f.history[1].energy == -0.7425606200908043e-01
    np.testing.assert_array_equal(f.history_energy, [h.energy for h in f.history])


def test_sorted_write(golden_dir):
    f = read_inout(_fixture(golden_dir, "inout_1_500"))
    K = f.basis_size
    perm = np.random.Generator(np.random.PCG64(7)).permutation(K) + 1
    lines = format_inout(f, func_num=perm).splitlines()
    plain = format_inout(f).splitlines()
    first = len(plain) - K
    assert lines[:first] == plain[:first]
    for k in range(K):       # function k goes to line perm[k]
        assert lines[first + perm[k] - 1][8:] == plain[first + k][8:]
    with pytest.raises(ValueError):
        format_inout(f, func_num=np.ones(K, int))


# line images of '(1x,a9,1x,a1,...)' formats, column by column
_D = " 0.1000000000000000D-05"        # D23.16 of 1e-6 (23 columns)
_R = " 0.2500000000000000D+01"        # D23.16 of 2.5
BBOP_LINES = [
    (BbopStep("FULL_OPT1", "I", [1, 500, 3, 40, 2, 1], [1e-6, 2.5], ["hess.dat"]),
     " FULL_OPT1 I      1    500      3     40 " + _D + " " + _R + "      2      1 hess.dat"),
    (BbopStep("EXPC_VALS", "G", [12]), " EXPC_VALS G     12"),
    (BbopStep("DENSITIES", "G", [3], [], ["a.txt", "b.txt", "c.txt", "d.txt"]),
     " DENSITIES G      3 a.txt b.txt c.txt d.txt"),
    (BbopStep("ELIM_LCFN", "G", [10], [1e-6], ["elim.txt"]), " ELIM_LCFN G     10" + _D + " elim.txt"),
    (BbopStep("ELIM_LND1", "I", [10], [-2.5], ["e.txt"]),
     " ELIM_LND1 I     10-0.2500000000000000D+01 e.txt"),
    (BbopStep("SEPR_LND1", "G", [4], [1e-6, 2.5], ["s.txt"]), " SEPR_LND1 G      4" + _D + " " + _R + " s.txt"),
    (BbopStep("SEPR_FLCF", "I", [4], [1e-6, 2.5], ["s.txt"]), " SEPR_FLCF I      4" + _D + " " + _R + " s.txt"),
    (BbopStep("SAVE_FILE", "", [100], [], ["save.txt"]), " SAVE_FILE    100 save.txt"),
]


@pytest.mark.parametrize("step,line", BBOP_LINES, ids=[s.action for s, _ in BBOP_LINES])
def test_bbop_line_formats(step, line):
    assert _format_bbop(step) == line
    assert _parse_bbop(line) == step


def test_fortran_d_cases():
    assert fortran_d(0.0) == " 0.0000000000000000D+00"
    assert fortran_d(-0.5978979662615874) == "-0.5978979662615874D+00"
    assert fortran_d(1e100) == " 0.1000000000000000+101"
    assert fortran_d(1e-100) == " 0.1000000000000000D-99"
    assert fortran_d(1e-300) == " 0.1000000000000000-299"
    assert fortran_d(0.09999999999999999) == " 0.9999999999999999D-01"
    assert fortran_d(123456789.123456789) == " 0.1234567891234568D+09"


_FD_PROGRAM = """program fd
  implicit none
  double precision :: x
  integer :: ios
  do
    read(*,*,iostat=ios) x
    if (ios /= 0) exit
    write(*,'(d23.16)') x
  end do
end program fd
"""


@pytest.mark.skipif(shutil.which("flang") is None, reason="no Fortran compiler to compare with")
def test_fortran_d_matches_fortran_compiler(tmp_path):
    src = tmp_path / "fd.f90"
    src.write_text(_FD_PROGRAM)
    exe = tmp_path / "fd"
    subprocess.run(["flang", "-O1", "-o", str(exe), str(src)], check=True, capture_output=True)
    rng = np.random.Generator(np.random.PCG64(11))
    vals = [0.0, 1.0, -1.0, 0.5, 1e100, -1e-100, 1e-300, 5e-324, -1.7976931348623157e308,
            0.9999999999999999, 0.09999999999999999, 9.9999999999999995e-8]
    vals += list(rng.uniform(-1, 1, 300) * 10.0 ** rng.integers(-40, 40, 300))
    vals += list(rng.normal(size=100))
    text = "\n".join(repr(float(v)) for v in vals) + "\n"
    got = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.splitlines()
    assert got == [fortran_d(float(v)) for v in vals]


# ------------------------------------------------------------------ swap file
def test_swap_byte_layout(tmp_path):
    """K = 2 by hand: int32 K, W column-major (H on/below the diagonal, S
    above), diagS."""
    H = np.zeros((2, 2), np.complex128, order="F")
    S = np.zeros((2, 2), np.complex128, order="F")
    H[1, 0] = 3 + 4j
    S[1, 0] = 0.25 - 0.5j
    path = tmp_path / "swap"
    write_swap(str(path), H, S, [1.5, 2.5], [1.0, 1.0], "G")
    raw = path.read_bytes()
    expect = np.int32(2).tobytes() + np.array([1.5, 0, 3, 4, 0.25, -0.5, 2.5, 0, 1, 1]).tobytes()
    assert raw == expect


def _small_matrices(golden_dir, method, sigma=0.0):
    f = read_inout(_fixture(golden_dir, "inout_1_500"))
    return oracle.assemble(f.system, f.params[:40], method, sigma)


def test_swap_round_trip_G(golden_dir, tmp_path):
    H, S, dH, dS = _small_matrices(golden_dir, "G")
    K = dS.shape[0]
    path = str(tmp_path / "swap")
    write_swap(path, H, S, dH, dS, "G")
    Hr, Sr, dHr, dSr = read_swap(path, K, "G")
    low = np.tril_indices(K, -1)
    np.testing.assert_array_equal(Hr[low], H[low])
    np.testing.assert_array_equal(Sr[low], S[low])
    np.testing.assert_array_equal(dHr, dH)
    np.testing.assert_array_equal(dSr, dS)
    np.testing.assert_array_equal(Hr.diagonal(), dH)      # as the reference restores Glob_H
    assert open(path, "rb").read() == EMPTY_MARK
    assert read_swap(path, K, "G") is None


def test_swap_round_trip_I_and_mode_independence(golden_dir, tmp_path):
    sigma = -0.59
    HG, SG, dHG, dSG = _small_matrices(golden_dir, "G")
    HI, SI, dHI, dSI = _small_matrices(golden_dir, "I", sigma)
    K = dSG.shape[0]
    pg, pi = str(tmp_path / "g"), str(tmp_path / "i")
    write_swap(pg, HG, SG, dHG, dSG, "G")
    write_swap(pi, HI, SI, dHI, dSI, "I", sigma)
    bg, bi = open(pg, "rb").read(), open(pi, "rb").read()
    # the file holds unshifted H whichever layout wrote it
    wg = np.frombuffer(bg[4:4 + 16 * K * K], np.complex128)
    wi = np.frombuffer(bi[4:4 + 16 * K * K], np.complex128)
    np.testing.assert_allclose(wi, wg, rtol=0, atol=1e-14 * np.abs(wg).max())
    assert bg[-8 * K:] == bi[-8 * K:]
    # a file written in G layout read back in I layout gives the I arrays
    Hr, Sr, _, _ = read_swap(pg, K, "I", sigma)
    np.testing.assert_array_equal(Sr, SI)
    low = np.tril_indices(K)
    np.testing.assert_allclose(Hr[low], HI[low], rtol=0, atol=1e-14 * np.abs(HI).max())


def test_swap_acceptance_rules(golden_dir, tmp_path):
    H, S, dH, dS = _small_matrices(golden_dir, "G")
    K = dS.shape[0]
    path = str(tmp_path / "swap")
    write_swap(path, H, S, dH, dS, "G")
    raw = open(path, "rb").read()
    assert read_swap(path, K + 1, "G") is None
    assert open(path, "rb").read() == raw          # a basis of another size leaves it alone
    with open(path, "wb") as fh:
        fh.write(raw[:-8])
    assert read_swap(path, K, "G") is None          # incomplete: rejected, and erased
    assert open(path, "rb").read() == EMPTY_MARK
    assert read_swap(str(tmp_path / "missing"), K, "G") is None
    with pytest.raises(ValueError):
        read_swap(path, K, "X")
