"""Seeded synthetic effective-Hamiltonian workloads built from this package's
own MPO builder (``ultradmrg_amd.mpo``), for the matvec benchmark.

An operator set is ``[L, W_l, W_r, R]`` at the middle bond of a chain or
cylinder MPO, with environment bonds of bond dimension ~D whose sectors follow
a Gaussian profile over the charges (U(1): 2Sz in -2K..2K; U(1)xU(1):
(N, 2Sz) in [-K, K]^2 with N + 2Sz even), and a state psi[a, s1, s2, b].  Every
structurally allowed block of L, R and psi is present and filled with
U[-1, 1) from ``numpy.random.default_rng(seed)`` in lexicographic block
order (L first, then R, then psi).
"""
from __future__ import annotations

import itertools
from typing import List, Tuple

import numpy as np
import torch

from . import mpo as M
from .tensor import IN, OUT, Axis, BlockTensor, _flow


def gaussian_bond(D: int, nsect_half: int, width: float, qn_len: int = 1):
    """Sector list [(qn, deg)] with deg ~ exp(-|q|^2 / (2 width^2)) scaled to a total of ~D."""
    if qn_len == 1:
        qs = [(2 * k,) for k in range(-nsect_half, nsect_half + 1)]
        w = np.array([np.exp(-(q[0] / 2) ** 2 / (2 * width ** 2)) for q in qs])
    else:
        qs = [(n, m) for n in range(-nsect_half, nsect_half + 1) for m in range(-nsect_half, nsect_half + 1)
              if (n + m) % 2 == 0]
        w = np.array([np.exp(-(q[0] ** 2 + q[1] ** 2) / (2 * width ** 2)) for q in qs])
    degs = np.maximum(1, np.round(w / w.sum() * D)).astype(int)
    return [(q, int(d)) for q, d in zip(qs, degs)]


def _allowed(axes: List[Axis], div) -> List[Tuple[int, ...]]:
    return [c for c in itertools.product(*[range(ax.nsect) for ax in axes]) if _flow(axes, c) == tuple(div)]


def _random_tensor(axes: List[Axis], div, rng, device, dtype) -> BlockTensor:
    blocks = {}
    for c in _allowed(axes, div):
        blocks[c] = rng.uniform(-1, 1, size=tuple(ax.degs[s] for ax, s in zip(axes, c)))
    return BlockTensor.from_blocks(axes, blocks, device=device, dtype=dtype)


def synthetic_eff_ham(kind: str, D: int, seed: int, nsect_half=None, width=None, device="cuda",
                      dtype=torch.float64):
    """Random L, W_l, W_r, R and psi at bond dimension ~D (see the module docstring).

    kind = "heisenberg": U(1) spin-1/2 chain MPO (w = 5) from an 8-site chain;
    kind = "hubbard": U(1)xU(1) Hubbard 4x16 snake cylinder MPO (t=1, U=8),
    middle bond (w = 18, d = 4)."""
    rng = np.random.default_rng(seed)
    if kind == "heisenberg":
        m = M.heisenberg_chain(8)
        mpo = M.build_mpo(m, device=device, dtype=dtype)
        Wl, Wr = mpo[3], mpo[4]
        nsect_half = nsect_half or 12
        width = width or max(1.0, nsect_half / 2.5)
        secs = gaussian_bond(D, nsect_half, width, 1)
    elif kind == "hubbard":
        m = M.hubbard_cylinder(16, 4)
        mpo = M.build_mpo(m, device=device, dtype=dtype)
        mid = m.n // 2
        Wl, Wr = mpo[mid - 1], mpo[mid]
        nsect_half = nsect_half or 5
        width = width or 2.0
        secs = gaussian_bond(D, nsect_half, width, 2)
    else:
        raise ValueError(f"unknown workload kind {kind!r}")
    zero = tuple(0 for _ in secs[0][0])
    a_out = Axis.make(secs, OUT)
    b_out = Axis.make(secs, OUT)
    pout = M.physical_axis(m.qns_at(0), OUT)
    L = _random_tensor([a_out.inverse(), Wl.axes[0].inverse(), a_out], zero, rng, device, dtype)
    R = _random_tensor([b_out, Wr.axes[3].inverse(), b_out.inverse()], zero, rng, device, dtype)
    psi = _random_tensor([a_out.inverse(), pout, pout, b_out], zero, rng, device, dtype)
    return [L, Wl, Wr, R], psi
