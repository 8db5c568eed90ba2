"""Inputs of the reference's Lanczos / expmv known-answer tests, rebuilt exactly.

``tests/testing_utils.h:26-54`` fills the random matrices with glibc ``rand()``
after ``srand(0)``; the same libc is called here through ctypes, so the
operators are bit-identical to the reference test's.  Shapes and index
directions follow ``tests/test_algorithm/test_lanczos_solver.cc:28-39,117-143``
and ``test_lanczos_expmv_solver.cc:33-35,114-122,279-304``.
"""
import ctypes

import numpy as np

from oracle.bsten import IN, OUT, BSTensor, Index

_libc = ctypes.CDLL("libc.so.6")
_libc.rand.restype = ctypes.c_int
RAND_MAX = 2147483647


def srand(seed):
    _libc.srand(ctypes.c_uint(seed))


def Rand():
    return _libc.rand() / RAND_MAX


def rand_real_sym_mat(dim):
    """``testing_utils.h:28-41``."""
    m = np.zeros((dim, dim))
    srand(0)
    for i in range(dim):
        for j in range(dim):
            m[i, j] = Rand()
    srand(0)
    for i in range(dim):
        for j in range(dim):
            m[j, i] += Rand()
    return m


def rand_cplx_her_mat(dim):
    """``testing_utils.h:43-54`` (continues the current rand() stream)."""
    m = np.zeros((dim, dim), dtype=np.complex128)
    for i in range(dim):
        for j in range(i):
            re = Rand()
            im = Rand()
            m[i, j] = re + 1j * im
            m[j, i] = re - 1j * im
    for i in range(dim):
        m[i, i] = Rand()
    return m


def trivial_indexes(d, D, dh):
    q0 = (0,)
    din = Index.make([(q0, d)], IN)
    Din = Index.make([(q0, D)], IN)
    vin = Index.make([(q0, dh)], IN)
    return din, din.inverse(), Din, Din.inverse(), vin, vin.inverse()


def kat_eff_ham(D, d, dh, complex_=False):
    """{L, Wl, Wr, R} of the reference KAT: L(i,k,j)=M[i,j], R(j,k,i)=M[i,j], W(k,i,j,k)=m[i,j]."""
    din, dout, Din, Dout, vin, vout = trivial_indexes(d, D, dh)
    dt = np.complex128 if complex_ else np.float64
    if complex_:
        M = rand_cplx_her_mat(D)
        msite = rand_cplx_her_mat(d)
    else:
        M = rand_real_sym_mat(D)
        msite = rand_real_sym_mat(d)
    L = BSTensor([Din, vout, Dout], dt)
    R = BSTensor([Dout, vin, Din], dt)
    W = BSTensor([vin, din, dout, vout], dt)
    Lb = np.zeros((D, dh, D), dt)
    Rb = np.zeros((D, dh, D), dt)
    Wb = np.zeros((dh, d, d, dh), dt)
    for i in range(D):
        for j in range(D):
            for k in range(dh):
                Lb[i, k, j] = M[i, j]
                Rb[j, k, i] = M[i, j]
    for i in range(d):
        for j in range(d):
            for k in range(dh):
                Wb[k, i, j, k] = msite[i, j]
    L.blocks[(0, 0, 0)] = Lb
    R.blocks[(0, 0, 0)] = Rb
    W.blocks[(0, 0, 0, 0)] = Wb
    return L, W, W.copy(), R, (Din, dout, Dout)
