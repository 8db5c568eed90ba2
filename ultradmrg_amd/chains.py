"""Independent random-field Heisenberg chains and their centre-site effective
Hamiltonians: the workload of the batching form (many chains solved in one
launch).

Chain c has L_c sites (uniform in [lmin, lmax]), J = 1, fields h_i ~ U(-2, 2)
and bond dimension D_c drawn from ``dims``, all from one seed.  Its state is a
seeded random MPS with total Sz = 0 (1/2 for odd L; raw site tensors, no
canonicalisation);
the environments of the centre two-site block are built by the GPU
environment updates (UpdateSiteLenvs / UpdateSiteRenvs,
``vmps_init.h:343-373``) from the chain ends, each rescaled to unit norm.
Such environments keep the effective Hamiltonian Hermitian (they are
A^dag W A contractions of a Hermitian MPO), so Lanczos applies.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from . import mpo as M
from .sweep import random_mps, trivial_env
from .tensor import BlockTensor, contract, env_update


@dataclass
class Chain:
    length: int
    D: int
    fields: np.ndarray
    eff_ham: List[BlockTensor]   # [L, W_l, W_r, R] at the centre bond
    psi: BlockTensor             # A_mid . A_mid+1


def _unit(t: BlockTensor) -> BlockTensor:
    nrm = torch.linalg.vector_norm(t.data)
    return BlockTensor(t.axes, t.coords, t.data / nrm, t.offsets)


def random_field_chains(n: int, seed: int, lmin: int = 32, lmax: int = 128, dims=(128, 256, 512, 768),
                        device="cuda") -> List[Chain]:
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        D = int(rng.choice(dims))
        h = rng.uniform(-2.0, 2.0, size=L)
        m = M.heisenberg_chain(L, fields=list(h))
        mpo = M.build_mpo(m, device=device)
        # total 2Sz = 0 for even L, 1 for odd L
        mps = random_mps(m, (L % 2,), D, seed=int(rng.integers(1 << 31)), device=device, canonical=False)
        mid = L // 2 - 1
        dt = mps[0].data.dtype
        le = trivial_env(mps[0].axes[0], mpo[0].axes[0], device, dt)
        for i in range(mid):
            le = _unit(env_update(le, mps[i], mpo[i], "l"))
        re = trivial_env(mps[-1].axes[2], mpo[-1].axes[3], device, dt)
        for i in range(L - 1, mid + 1, -1):
            re = _unit(env_update(re, mps[i], mpo[i], "r"))
        psi = _unit(contract(mps[mid], mps[mid + 1], 1))
        out.append(Chain(L, D, h, [le, mpo[mid], mpo[mid + 1], re], psi))
    return out
