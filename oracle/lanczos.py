"""Oracle: restatement of the reference's Krylov solvers and effective-Hamiltonian
matvecs, on top of ``oracle.bsten``.

TEST INFRASTRUCTURE ONLY (see ``oracle/bsten.py``).

Functions and the reference lines they follow:

* ``eff_ham_mul_two_site_state``   ``include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:260-289``
* ``eff_ham_mul_single_site_state`` ``lanczos_vmps_solver_impl.h:291-304``
* ``LanczosSolver``                ``lanczos_vmps_solver_impl.h:114-248``
* ``TridiagGsSolver``              ``lanczos_vmps_solver_impl.h:306-355`` (LAPACK ``dstev``:
  called here through ``scipy.linalg.eigh_tridiagonal(lapack_driver='stev')``,
  which is the same LAPACK routine)
* ``LanczosExpmvSolver``           ``include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h:70-184``
* ``TridiagExpme1Solver``          ``lanczos_expmv_solver_impl.h:199-240``
* ``Distance``                     ``lanczos_expmv_solver_impl.h:252-269``

Parity status: ``TridiagExpme1Solver`` is pinned by the reference's MATLAB
golden vectors (``tests/test_algorithm/test_lanczos_expmv_solver.cc:70-109``);
``LanczosSolver`` and both matvecs by the dense-``dsyev`` known-answer test of
``tests/test_algorithm/test_lanczos_solver.cc:42-167`` and block-sparse vs
dense brute force; ``LanczosExpmvSolver`` by the dense ``exp(-i tau H) v``
check of ``test_lanczos_expmv_solver.cc:192-272``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Sequence

import numpy as np
import scipy.linalg as sla

from .bsten import BSTensor, Contract, ContractRotated, LinearCombine, vdot


@dataclass
class LanczosParams:
    """``include/qlmps/algorithm/lanczos_params.h:27-46`` (defaults 1e-7 / 200)."""
    error: float = 1.0e-7
    max_iterations: int = 200


@dataclass
class LanczosRes:
    iters: int
    gs_eng: float
    gs_vec: BSTensor


@dataclass
class ExpmvRes:
    iters: int
    expmv: BSTensor


# --------------------------------------------------------------------------
# matvecs
# --------------------------------------------------------------------------
def eff_ham_mul_two_site_state(eff_ham: Sequence[BSTensor], state: BSTensor) -> BSTensor:
    """Four chained contractions, ``lanczos_vmps_solver_impl.h:268-272``.

    L[a',w,a] psi[a,s1,s2,b] -> T1[a',w,s1,s2,b]
    T1 x Wl over (w,s1)      -> T2[s2,b,a',s1',w1]
    T2 x Wr over (w1,s2)     -> T3[b,a',s1',s2',w2]
    T3 x R  over (w2,b)      -> res[a',s1',s2',b']
    """
    L, Wl, Wr, R = eff_ham
    t1 = Contract(L, state, ([2], [0]))
    t2 = ContractRotated(t1, Wl, 1, 0, 2)
    t3 = ContractRotated(t2, Wr, 4, 0, 2)
    return ContractRotated(t3, R, 4, 1, 2)


def eff_ham_mul_single_site_state(eff_ham: Sequence[BSTensor], state: BSTensor) -> BSTensor:
    """Three chained contractions, ``lanczos_vmps_solver_impl.h:299-302``."""
    L, W, R = eff_ham
    t1 = Contract(L, state, ([2], [0]))
    t2 = ContractRotated(t1, W, 1, 0, 2)
    return ContractRotated(t2, R, 3, 1, 2)


def two_site_matvec_flops(eff_ham: Sequence[BSTensor], state: BSTensor) -> int:
    """Sum of 2*M*N*K (8*M*N*K complex) over the block pairs of the four steps."""
    L, Wl, Wr, R = eff_ham
    total = 0
    cplx = np.iscomplexobj(L.blocks and next(iter(L.blocks.values()))) or state.dtype.kind == "c"

    def pair_flops(A, B, axA, axB):
        nonlocal total
        bucket = {}
        for kb, vb in B.blocks.items():
            bucket.setdefault(tuple(kb[i] for i in axB), []).append(vb)
        for ka, va in A.blocks.items():
            key = tuple(ka[i] for i in axA)
            kdim = int(np.prod([va.shape[i] for i in axA]))
            m = va.size // max(kdim, 1)
            for vb in bucket.get(key, ()):
                n = vb.size // max(kdim, 1)
                total += (8 if cplx else 2) * m * n * kdim

    t1 = Contract(L, state, ([2], [0]))
    pair_flops(L, state, [2], [0])
    pair_flops(t1, Wl, [1, 2], [0, 1])
    t2 = ContractRotated(t1, Wl, 1, 0, 2)
    pair_flops(t2, Wr, [4, 0], [0, 1])
    t3 = ContractRotated(t2, Wr, 4, 0, 2)
    pair_flops(t3, R, [4, 0], [1, 2])
    return total


# --------------------------------------------------------------------------
# tridiagonal helpers
# --------------------------------------------------------------------------
def TridiagGsSolver(a: Sequence[float], b: Sequence[float], n: int, jobz: str):
    """Lowest eigenpair of the symmetric tridiagonal T(a[:n], b[:n-1]) via LAPACK dstev.

    Returns ``(gs_eng, gs_vec)``; ``gs_vec`` is None for ``jobz == 'N'``.
    ``lanczos_vmps_solver_impl.h:306-355``.
    """
    d = np.asarray(a[:n], dtype=np.float64)
    e = np.asarray(b[:n - 1], dtype=np.float64)
    if n == 1:
        return float(d[0]), (np.ones(1) if jobz == "V" else None)
    if jobz == "N":
        w = sla.eigh_tridiagonal(d, e, eigvals_only=True, lapack_driver="stev")
        return float(w[0]), None
    w, z = sla.eigh_tridiagonal(d, e, eigvals_only=False, lapack_driver="stev")
    return float(w[0]), z[:, 0].copy()


def TridiagExpme1Solver(a: Sequence[float], b: Sequence[float], n: int, step_length: float) -> np.ndarray:
    """exp(-i*step*T) e_1 = Z diag(exp(-i*step*lambda)) Z^T e_1, ``lanczos_expmv_solver_impl.h:199-240``."""
    d = np.asarray(a[:n], dtype=np.float64)
    if n == 1:
        return np.array([np.exp(-1j * step_length * d[0])])
    e = np.asarray(b[:n - 1], dtype=np.float64)
    w, z = sla.eigh_tridiagonal(d, e, eigvals_only=False, lapack_driver="stev")
    t = np.exp(-1j * step_length * w) * z[0, :]
    return z @ t


def Distance(v1: np.ndarray, v2: np.ndarray, n: int) -> float:
    """||v1 - v2||_2 over the first n entries, ``lanczos_expmv_solver_impl.h:252-269``."""
    return float(np.linalg.norm(np.asarray(v1[:n]) - np.asarray(v2[:n])))


# --------------------------------------------------------------------------
# Lanczos ground state
# --------------------------------------------------------------------------
def _linear_combine_bases(coefs, bases: List[BSTensor], proto: BSTensor) -> BSTensor:
    res = BSTensor(proto.indexes, proto.dtype)
    return LinearCombine(list(coefs), bases, 0.0, res)


def LanczosSolver(eff_ham, init_state: BSTensor, matvec: Callable, params: LanczosParams,
                  trace: list | None = None) -> LanczosRes:
    """Plain three-term Lanczos with no reorthogonalisation, ``lanczos_vmps_solver_impl.h:114-248``.

    Stop when ``(E_old - E_new) < error`` (absolute, signed), or the Krylov
    dimension reaches ``init_state.size()``, or ``m == max_iterations - 1``.
    ``trace`` (optional) receives ``(a, b)`` coefficient lists for testing.
    """
    dim = init_state.size()
    maxit = params.max_iterations
    bases: List[BSTensor] = [None] * maxit
    a = [0.0] * maxit
    b = [0.0] * maxit
    N = [0.0] * maxit
    state = init_state.copy()
    state.normalize()
    bases[0] = state
    last = matvec(eff_ham, bases[0])
    a[0] = float(np.real(vdot(last, bases[0])))
    N[0] = 0.0
    m = 0
    energy0 = a[0]
    while True:
        m += 1
        gamma = last
        if m == 1:
            LinearCombine([-a[m - 1]], [bases[m - 1]], 1.0, gamma)
        else:
            LinearCombine([-a[m - 1], -np.sqrt(N[m - 1])], [bases[m - 1], bases[m - 2]], 1.0, gamma)
        norm_gamma = gamma.normalize()
        if norm_gamma == 0.0:
            if m == 1:
                return LanczosRes(m, energy0, bases[0].copy())
            _, vec = TridiagGsSolver(a, b, m, "V")
            gs = _linear_combine_bases(vec, bases[:m], bases[0])
            return LanczosRes(m, energy0, gs)
        N[m] = norm_gamma * norm_gamma
        b[m - 1] = norm_gamma
        bases[m] = gamma
        last = matvec(eff_ham, bases[m])
        a[m] = float(np.real(vdot(last, bases[m])))
        energy0_new, _ = TridiagGsSolver(a, b, m + 1, "N")
        if (energy0 - energy0_new) < params.error or m == dim or m == maxit - 1:
            energy0_new, vec = TridiagGsSolver(a, b, m + 1, "V")
            energy0 = energy0_new
            gs = _linear_combine_bases(vec, bases[:m + 1], bases[0])
            if trace is not None:
                trace.extend([a[:m + 1], b[:m]])
            return LanczosRes(m + 1, energy0, gs)
        energy0 = energy0_new


# --------------------------------------------------------------------------
# Lanczos expmv (TDVP)
# --------------------------------------------------------------------------
def LanczosExpmvSolver(eff_ham, init_state: BSTensor, matvec: Callable, step_length: float,
                       params: LanczosParams) -> ExpmvRes:
    """exp(-i*step*H_eff) psi by Lanczos, ``lanczos_expmv_solver_impl.h:70-184``.

    Keeps ||psi0|| (``Normalize`` returns it and the result is rescaled);
    converged when ``Distance(c_prev, c) < error`` (c_prev starts at zero), or
    the Krylov dimension reaches ``size()``, or ``m == max_iterations - 1``.
    """
    dim = init_state.size()
    maxit = params.max_iterations
    bases: List[BSTensor] = [None] * maxit
    a = [0.0] * maxit
    b = [0.0] * maxit
    state = init_state.copy()
    initial_norm = state.normalize()
    bases[0] = state
    last = matvec(eff_ham, bases[0])
    a[0] = float(np.real(vdot(last, bases[0])))
    m = 0
    last_c = np.zeros(maxit, dtype=np.complex128)
    while True:
        m += 1
        gamma = last
        if m == 1:
            LinearCombine([-a[m - 1]], [bases[m - 1]], 1.0, gamma)
        else:
            LinearCombine([-a[m - 1], -b[m - 2]], [bases[m - 1], bases[m - 2]], 1.0, gamma)
        norm_gamma = gamma.normalize()
        if norm_gamma == 0.0:
            if m == 1:
                out = bases[0].copy().scale(initial_norm * np.exp(-1j * step_length * a[0]))
                return ExpmvRes(m, out)
            c = TridiagExpme1Solver(a, b, m, step_length) * initial_norm
            return ExpmvRes(m, _linear_combine_bases(c, bases[:m], bases[0]))
        b[m - 1] = norm_gamma
        bases[m] = gamma
        last = matvec(eff_ham, bases[m])
        a[m] = float(np.real(vdot(last, bases[m])))
        c = np.zeros(maxit, dtype=np.complex128)
        c[:m + 1] = TridiagExpme1Solver(a, b, m + 1, step_length)
        dist = Distance(last_c, c, m + 1)
        if dist < params.error or m == dim or m == maxit - 1:
            cc = c[:m + 1] * initial_norm
            return ExpmvRes(m + 1, _linear_combine_bases(cc, bases[:m + 1], bases[0]))
        last_c = c
