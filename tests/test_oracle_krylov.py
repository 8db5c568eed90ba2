"""Oracle Krylov solvers against the reference's own known-answer tests."""
import numpy as np
import pytest

from oracle.bsten import BSTensor, random_fill
from oracle.lanczos import (LanczosExpmvSolver, LanczosParams, LanczosSolver, TridiagExpme1Solver,
                            TridiagGsSolver, Distance, eff_ham_mul_single_site_state,
                            eff_ham_mul_two_site_state)
from tests import refkat

# MATLAB golden vectors, tests/test_algorithm/test_lanczos_expmv_solver.cc:70-91
GOLD_2 = dict(a=[0.5, 0.3], b=[0.2], delta=-1.3, res=[
    0.76772272947713149360282613997697 + 0.58726872368826332770908038583002j,
    -0.12737709795879115226568956131814 + 0.22246872080662932757988414778083j])
GOLD_5 = dict(a=[1.8, 2.4, 0.5, 6.3, 0.3], b=[1.1, 0.2, 8.5, 0.9], delta=1.5, res=[
    0.13627869919086488259551970259054 - 0.2627770725485846226021635629877j,
    0.0095890697466718029240428933235307 + 0.9546547280068733432045746667427j,
    -0.000013884256143736063197576847960679 + 0.013856219332059563353887199355086j,
    0.0018295388155743035323713696627124 - 0.022002544240195875957510907028336j,
    -0.0094873011859597823625112056333819 - 0.012217334778703219794193657321557j])


@pytest.mark.parametrize("g", [GOLD_2, GOLD_5])
def test_tridiag_expme1_matlab_golden(g):
    n = len(g["a"])
    res = TridiagExpme1Solver(g["a"], g["b"], n, g["delta"])
    assert Distance(res, np.array(g["res"]), n) < 1e-13


def test_tridiag_gs_matches_dense():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 40, 100):
        a, b = rng.normal(size=n), rng.normal(size=max(n - 1, 0))
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        w, V = np.linalg.eigh(T)
        e, v = TridiagGsSolver(list(a), list(b), n, "V")
        assert abs(e - w[0]) < 1e-12 * max(1, abs(w[0]))
        assert abs(abs(v @ V[:, 0]) - 1) < 1e-10


def _dense_two_site(L, Wl, Wr, R):
    Ld, Wld, Wrd, Rd = (t.to_dense() for t in (L, Wl, Wr, R))
    H = np.einsum("xwa,wsty,yuvz,bzc->xtvbasuc", Ld, Wld, Wrd, Rd, optimize=True)
    n = Ld.shape[0] * Wld.shape[2] * Wrd.shape[2] * Rd.shape[0]
    return H.reshape(n, n)


def _dense_one_site(L, W, R):
    Ld, Wd, Rd = (t.to_dense() for t in (L, W, R))
    H = np.einsum("xwa,wsty,byc->xtbasc", Ld, Wd, Rd, optimize=True)
    n = Ld.shape[0] * Wd.shape[2] * Rd.shape[0]
    return H.reshape(n, n)


@pytest.mark.parametrize("params", [LanczosParams(1e-9), LanczosParams(1e-16, 20)])
def test_lanczos_two_site_kat_real(params):
    """test_lanczos_solver.cc:114-167: D=20, d=2, dh=2, |E - dsyev| < 1e-8, both stop rules."""
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2)
    psi = random_fill(BSTensor([Din, dout, dout, Dout]), (0,), np.random.default_rng(0))
    res = LanczosSolver([L, Wl, Wr, R], psi, eff_ham_mul_two_site_state, params)
    w0 = np.linalg.eigvalsh(_dense_two_site(L, Wl, Wr, R))[0]
    assert abs(res.gs_eng - w0) < 1e-8
    if params.max_iterations == 20:
        assert res.iters == 20


def test_lanczos_two_site_kat_complex():
    refkat.srand(0)
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2, complex_=True)
    psi = random_fill(BSTensor([Din, dout, dout, Dout], np.complex128), (0,), np.random.default_rng(1), True)
    res = LanczosSolver([L, Wl, Wr, R], psi, eff_ham_mul_two_site_state, LanczosParams(1e-9))
    H = _dense_two_site(L, Wl, Wr, R)
    assert np.allclose(H, H.conj().T)
    assert abs(res.gs_eng - np.linalg.eigvalsh(H)[0]) < 1e-8


def test_lanczos_single_site_kat():
    """test_lanczos_solver.cc:282-324."""
    L, W, _, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2)
    psi = random_fill(BSTensor([Din, dout, Dout]), (0,), np.random.default_rng(2))
    res = LanczosSolver([L, W, R], psi, eff_ham_mul_single_site_state, LanczosParams(1e-9))
    assert abs(res.gs_eng - np.linalg.eigvalsh(_dense_one_site(L, W, R))[0]) < 1e-8


def test_lanczos_expmv_kat():
    """test_lanczos_expmv_solver.cc:221-315: D=10, tau=0.1, LanczosParams(1e-14, 100), rel. err < 1e-13."""
    refkat.srand(0)
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(10, 2, 2, complex_=True)
    psi = random_fill(BSTensor([Din, dout, dout, Dout], np.complex128), (0,), np.random.default_rng(4), True)
    v = psi.to_dense().reshape(-1)
    H = _dense_two_site(L, Wl, Wr, R)
    w, V = np.linalg.eigh(H)
    ref = V @ (np.exp(-1j * 0.1 * w) * (V.conj().T @ v))
    res = LanczosExpmvSolver([L, Wl, Wr, R], psi, eff_ham_mul_two_site_state, 0.1, LanczosParams(1e-14, 100))
    out = res.expmv.to_dense().reshape(-1)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-13
