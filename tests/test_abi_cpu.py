"""CPU-side checks of the C-ABI library: it loads, exports every declared
entry point, and its host FP64 tridiagonal helpers match the reference's
golden vectors and LAPACK (through scipy)."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.linalg as sla

from ultradmrg_amd import _lib
from ultradmrg_amd.solver import TridiagExpme1Solver, TridiagGsSolver
from tests.test_oracle_krylov import GOLD_2, GOLD_5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "qlgpu.h")).read()
    return sorted(set(re.findall(r"\b(qlg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_entry_points()
    assert names == sorted(_lib.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("g", [GOLD_2, GOLD_5])
def test_tridiag_expme1_matlab_golden(g):
    n = len(g["a"])
    res = TridiagExpme1Solver(g["a"], g["b"], n, g["delta"])
    assert np.linalg.norm(res - np.array(g["res"])) < 1e-13


def test_tridiag_gs_vs_lapack_stev():
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 10, 57, 100):
        a, b = rng.normal(size=n), rng.normal(size=max(n - 1, 0))
        e, v = TridiagGsSolver(list(a), list(b), n, "V")
        if n == 1:
            assert e == a[0]
            continue
        w, z = sla.eigh_tridiagonal(a, b, lapack_driver="stev")
        assert abs(e - w[0]) <= 4e-15 * np.abs(w).max()
        assert abs(abs(v @ z[:, 0]) - 1.0) < 1e-12
        e2, _ = TridiagGsSolver(list(a), list(b), n, "N")
        assert e2 == e


def test_tridiag_degenerate_and_zero_offdiag():
    # b = 0 splits the matrix: eigenvalues are the diagonal
    e, v = TridiagGsSolver([3.0, -1.0, 2.0], [0.0, 0.0], 3, "V")
    assert e == -1.0 and abs(abs(v[1]) - 1.0) < 1e-15


def test_invalid_arguments_return_status_not_exit():
    L = _lib.lib()
    a = (ctypes.c_double * 2)(1.0, 2.0)
    e = ctypes.c_double()
    assert L.qlg_tridiag_gs(a, a, 0, b"N", ctypes.byref(e), None) == 1
    assert L.qlg_tridiag_gs(a, a, 2, b"X", ctypes.byref(e), None) == 1
    assert L.qlg_matvec(None, None, None) == 1
