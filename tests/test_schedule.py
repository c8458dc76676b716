"""CPU check of the wait rule of the blocked LDL^H's trailing-update
launches (and the first panel's launch): the task maps the kernels decode (csrc/ecg_schedule.hpp) are
enumerated for every basis size up to 4096 and every launch; each awaited
flag must be published in the same launch by a workgroup of lower index,
and the tasks must cover the launch's work exactly once
(tests/sched_check.cpp).  A planted wait on a higher index must be caught."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path_factory.mktemp("sched") / "sched_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "sched_check.cpp")], check=True)
    return exe


def test_waits_point_to_lower_workgroups(checker):
    out = subprocess.run([checker, "4096"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("OK ")
    assert int(out.stdout.split()[1]) > 100000   # launches checked


def test_planted_higher_wait_is_caught(checker):
    out = subprocess.run([checker, "200", "plant"], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode != 0 and "not lower" in out.stdout


def test_batched_launches_wait_on_lower_linear_ids(checker):
    """The batched factorisation's launches (problem-major 2-D grids, uneven
    basis sizes): every wait names the same problem's flag at a lower
    linear workgroup index."""
    out = subprocess.run([checker, "batch", "300"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("OK ")
    assert int(out.stdout.split()[1]) > 1000


def test_batched_planted_cross_problem_wait_is_caught(checker):
    out = subprocess.run([checker, "batch", "5", "plant"], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode != 0 and "not its own" in out.stdout
