"""Block SVD host logic, on the CPU.

``sweep.block_svd`` (the reference's two-site SVD,
two_site_update_finite_vmps_impl.h:302) batches the sector matrices, hands
their Gram matrices to the GPU eigensolver, and then does everything else
with array operations: singular values from the back-projected rows,
Newton-Schulz orthogonalisation, the pooled truncation, and the cutting of
u, s and vt back into blocks.  These tests run that logic on the CPU with
torch's eigensolver standing in for the GPU call, and check the result
against numpy's dense SVD: exact reconstruction without truncation,
orthonormal u and vt, the kept singular values, and the truncation error."""
import itertools
import random

import numpy as np
import pytest
import torch

from ultradmrg_amd import sweep
from ultradmrg_amd.tensor import IN, OUT, Axis, BlockTensor


@pytest.fixture(autouse=True)
def cpu_eigensolver(monkeypatch):
    monkeypatch.setattr(sweep, "_eigh_grams", lambda grams: {b: torch.linalg.eigh(G) for b, G in grams.items()})


def _random_tensor(seed, cplx, deficient=False):
    rng = random.Random(seed)
    rank = rng.randint(3, 4)
    axes = []
    for _ in range(rank):
        qs = rng.sample(range(-2, 3), rng.randint(1, 3))
        axes.append(Axis.make([((q,), rng.randint(1, 7)) for q in qs], rng.choice([IN, OUT])))
    # blocks of one divergence (flow), as in any tensor of the sweep
    allc = list(itertools.product(*[range(ax.nsect) for ax in axes]))
    div = _flow(axes, rng.choice(allc))
    allc = [c for c in allc if _flow(axes, c) == div]
    coords = sorted(rng.sample(allc, rng.randint(1, min(len(allc), 20))))
    n = sum(int(np.prod([ax.degs[s] for ax, s in zip(axes, c)])) for c in coords)
    g = np.random.default_rng(seed)
    x = g.standard_normal(n) + (1j * g.standard_normal(n) if cplx else 0)
    if deficient:
        x = x * (g.random(n) < 0.1)
    return BlockTensor(axes, coords, torch.from_numpy(np.asarray(x))), rng.randint(1, rank - 1), div


def _flow(axes, key):
    return sum((ax.qns[s][0] if ax.dir == OUT else -ax.qns[s][0]) for ax, s in zip(axes, key))


def _dense(t):
    starts = [np.concatenate([[0], np.cumsum(ax.degs)[:-1]]).astype(int) for ax in t.axes]
    out = np.zeros([ax.dim for ax in t.axes], dtype=complex if t.is_complex else float)
    for c, blk in t.blocks().items():
        out[tuple(slice(o[s], o[s] + ax.degs[s]) for o, ax, s in zip(starts, t.axes, c))] = blk
    return out


def _sector_singular_values(t, ldims, div):
    """All singular values of the tensor's sector matrices, from numpy's dense SVD:
    a sector pairs the left keys of one flow f with the right keys of flow div - f."""
    d = _dense(t)
    left, right = t.axes[:ldims], t.axes[ldims:]
    lidx = list(itertools.product(*[range(ax.nsect) for ax in left]))
    ridx = list(itertools.product(*[range(ax.nsect) for ax in right]))
    starts = [np.concatenate([[0], np.cumsum(ax.degs)[:-1]]).astype(int) for ax in t.axes]

    def span(axs, st, key):
        return [np.arange(o[s], o[s] + ax.degs[s]) for o, ax, s in zip(st, axs, key)]

    svs = []
    for f in sorted({_flow(left, k) for k in lidx}):
        rows = [k for k in lidx if _flow(left, k) == f]
        cols = [k for k in ridx if _flow(right, k) == div - f]
        if not cols:
            continue
        M = np.block([[d[np.ix_(*span(left, starts[:ldims], r), *span(right, starts[ldims:], c))]
                       .reshape(int(np.prod([ax.degs[s] for ax, s in zip(left, r)])), -1) for c in cols] for r in rows])
        svs.append(np.linalg.svd(M, compute_uv=False))
    return np.sort(np.concatenate(svs))[::-1] if svs else np.zeros(0)


def _contract(res, ldims, t):
    u, s, vt = _dense(res.u), _dense(res.s), _dense(res.vt)
    lshape, rshape = u.shape[:-1], vt.shape[1:]
    U, V = u.reshape(-1, u.shape[-1]), vt.reshape(vt.shape[0], -1)
    return U, np.diag(s) if s.ndim == 1 else s, V, (U @ s @ V).reshape(lshape + rshape)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("seed", range(8))
def test_untruncated_svd_reconstructs(seed, cplx):
    t, ldims, _ = _random_tensor(seed, cplx)
    res = sweep.block_svd(t, ldims, (0,), 0.0, 1, 10**6)
    U, S, V, rec = _contract(res, ldims, t)
    ref = _dense(t)
    assert np.abs(rec - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(U.conj().T @ U, np.eye(U.shape[1]), atol=1e-12)
    np.testing.assert_allclose(V @ V.conj().T, np.eye(V.shape[0]), atol=1e-12)
    assert res.trunc_err == 0.0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("seed", range(8))
def test_truncation_keeps_largest_values(seed, cplx):
    t, ldims, div = _random_tensor(100 + seed, cplx)
    ref_sv = _sector_singular_values(t, ldims, div)
    for trunc in (1e-3, 1e-1):
        res = sweep.block_svd(t, ldims, (0,), trunc, 1, 10**6)
        U, S, V, rec = _contract(res, ldims, t)
        kept = np.sort(np.diag(S).real)[::-1]
        D = len(kept)
        assert D == res.D
        np.testing.assert_allclose(kept, ref_sv[:D], rtol=1e-10, atol=1e-12 * ref_sv[0])
        total = float((ref_sv ** 2).sum())
        discarded = float((ref_sv[D:] ** 2).sum()) / total
        assert discarded <= trunc + 1e-12
        if D > 1:   # one value fewer would discard too much
            assert float((ref_sv[D - 1:] ** 2).sum()) / total > trunc
        assert abs(res.trunc_err - discarded) <= 1e-10
        ref = _dense(t)
        assert abs(np.linalg.norm(rec - ref) ** 2 / total - discarded) <= 1e-10


@pytest.mark.parametrize("seed", range(6))
def test_kept_zero_singular_values_get_orthonormal_vectors(seed):
    t, ldims, _ = _random_tensor(200 + seed, seed % 2 == 1, deficient=True)
    res = sweep.block_svd(t, ldims, (0,), 0.0, 10**6, 10**6)
    U, S, V, rec = _contract(res, ldims, t)
    ref = _dense(t)
    assert np.abs(rec - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(U.conj().T @ U, np.eye(U.shape[1]), atol=1e-10)
    np.testing.assert_allclose(V @ V.conj().T, np.eye(V.shape[0]), atol=1e-10)
