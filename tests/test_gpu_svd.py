"""The block SVD in the library (qlg_block_svd, the reference's SVD call of
two_site_update_finite_vmps_impl.h:302-307) against numpy's dense SVD of each
sector matrix.

Sector matrices are built with prescribed singular values spanning 14 decades
(down to 1e-13 of the sector's largest), exact zeros (rank deficiency) and
both orientations (more rows than columns and the reverse), real and complex.
Bars: every kept singular value equals numpy's to 1e-12 of the tensor's
largest (no value is rounded to zero); the kept count and discarded weight
follow the pooled truncation rule exactly; u and vt have orthonormal columns /
rows per sector to 1e-12; u s vt reconstructs the tensor up to the discarded
weight."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.sweep import block_svd
    from ultradmrg_amd.tensor import Axis, BlockTensor, contract


def _orth(rng, k, cols, cplx):
    z = rng.standard_normal((k, cols)) + (1j * rng.standard_normal((k, cols)) if cplx else 0)
    return np.linalg.qr(z)[0]


def _case(cplx, seed=11):
    """rank-2 tensor [a (IN), b (OUT)] with sector matrices of prescribed spectra"""
    rng = np.random.default_rng(seed)
    dims = [(40, 25), (25, 40), (60, 60), (1, 7), (30, 12)]
    a_ax = Axis.make([((2 * i,), r) for i, (r, _) in enumerate(dims)], -1)
    b_ax = Axis.make([((2 * i,), c) for i, (_, c) in enumerate(dims)], 1)
    blocks, spectra = {}, {}
    for i, (r, c) in enumerate(dims):
        n = min(r, c)
        sv = np.logspace(0, -13, n) * (3.0 ** -i)
        if i == 4:
            sv[-3:] = 0.0               # exact rank deficiency
        blocks[(i, i)] = (_orth(rng, r, n, cplx) * sv) @ _orth(rng, c, n, cplx).conj().T
        spectra[i] = np.linalg.svd(blocks[(i, i)], compute_uv=False)
    return a_ax, b_ax, blocks, spectra


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("trunc,Dmin,Dmax", [(0.0, 10 ** 6, 10 ** 6), (1e-10, 1, 10 ** 6), (0.0, 20, 20)])
def test_block_svd_vs_numpy(cplx, trunc, Dmin, Dmax):
    print(f"[case] library block svd cplx={cplx} trunc={trunc} D=[{Dmin},{Dmax}]", flush=True)
    a_ax, b_ax, blocks, spectra = _case(cplx)
    t = BlockTensor.from_blocks([a_ax, b_ax], blocks)
    r = block_svd(t, 1, (0,), trunc, Dmin, Dmax)
    smax = max(float(s.max()) for s in spectra.values())
    # pooled truncation on numpy's values
    allv = np.sort(np.concatenate(list(spectra.values())))[::-1]
    tail = np.concatenate([np.cumsum((allv ** 2)[::-1])[::-1], [0.0]])
    D = next(k for k in range(len(allv) + 1) if tail[k] / tail[0] <= trunc)
    D = max(min(max(D, min(Dmin, len(allv))), Dmax, len(allv)), 1)
    assert r.D == D
    assert abs(r.trunc_err - tail[D] / tail[0]) <= 1e-12
    # kept values per sector equal numpy's largest ones
    sblk = r.s.blocks()
    mid = r.u.axes[-1]
    got_all = []
    for (mi, _), blk in sblk.items():
        q = mid.qns[mi]
        i = q[0] // 2
        got = np.real(np.diag(blk))
        got_all.append(got)
        ref = spectra[i][:len(got)]
        assert np.abs(got - ref).max() <= 1e-12 * smax, (i, np.abs(got - ref).max() / smax)
    assert sum(len(g) for g in got_all) == D
    # orthonormality per kept sector
    for (i, mi), U in r.u.blocks().items():
        assert np.abs(U.conj().T @ U - np.eye(U.shape[1])).max() < 1e-12
    for (mi, j), V in r.vt.blocks().items():
        assert np.abs(V @ V.conj().T - np.eye(V.shape[0])).max() < 1e-12
    # reconstruction error = discarded weight
    rec = contract(contract(r.u, r.s, 1), r.vt, 1).blocks()
    num = sum(np.sum(np.abs(rec.get(k, 0) - b) ** 2) for k, b in blocks.items())
    den = sum(np.sum(np.abs(b) ** 2) for b in blocks.values())
    assert abs(num / den - r.trunc_err) <= 1e-11


def test_block_svd_rank4_two_site_shape():
    """A two-site state psi[a, s1, s2, b] split after two axes (the sweep's call):
    kept values vs numpy on the assembled sector matrices, reconstruction."""
    from oracle.dmrg import block_svd as o_block_svd
    from oracle.bsten import BSTensor, Index, random_fill
    from tests.convert import to_device, to_oracle
    print("[case] library block svd two-site shape", flush=True)
    rng = np.random.default_rng(3)
    a = Index.make([((-2,), 3), ((0,), 7), ((2,), 4)], -1)
    s = Index.make([((1,), 1), ((-1,), 1)], 1)
    b = Index.make([((-2,), 5), ((0,), 6), ((2,), 2)], 1)
    psi = random_fill(BSTensor([a, s, s, b]), (0,), rng)
    ref = o_block_svd(psi, 2, (0,), 1e-12, 1, 10 ** 6)
    out = block_svd(to_device(psi), 2, (0,), 1e-12, 1, 10 ** 6)
    assert out.D == ref.D
    rs = sorted(np.concatenate([np.diag(v) for v in ref.s.blocks.values()]))[::-1]
    gs = sorted(np.concatenate([np.diag(v) for v in out.s.blocks().values()]))[::-1]
    assert np.abs(np.array(rs) - np.array(gs)).max() <= 1e-12 * rs[0]


@pytest.mark.parametrize("cplx", [False, True])
def test_block_svd_degenerate_gram(cplx):
    """Sectors whose Gram matrix a unitary rotation leaves diagonal: exactly
    orthonormal rows with the order a multiple of 16 (Gram = I, no padding),
    a scaled copy of them, and an all-zero sector, beside a generic one (the
    complex eigensolver returns NaN vectors on such input; the library then
    finishes those sectors by Jacobi rotations).  Values vs numpy, orthonormal
    u and vt, reconstruction (no truncation) to 1e-10 of the largest entry:
    the Newton-Schulz orthonormalisation moves u s vt by up to
    ~eps s_max^2 / s_min, 2e-10 for the generic sector's 1e-6 spread."""
    print(f"[case] library block svd degenerate Gram cplx={cplx}", flush=True)
    rng = np.random.default_rng(5)
    ph = np.exp(1j * 0.3) if cplx else 1.0
    dims = [(16, 40), (32, 48), (16, 16), (24, 30)]
    a_ax = Axis.make([((2 * i,), r) for i, (r, _) in enumerate(dims)], -1)
    b_ax = Axis.make([((2 * i,), c) for i, (_, c) in enumerate(dims)], 1)
    blocks = {}
    blocks[(0, 0)] = np.hstack([np.eye(16), np.zeros((16, 24))]) * ph
    blocks[(1, 1)] = 0.5 * np.hstack([np.zeros((32, 16)), np.eye(32)]) * ph
    blocks[(2, 2)] = np.zeros((16, 16)) * ph
    blocks[(3, 3)] = (_orth(rng, 24, 24, cplx) * np.logspace(0, -6, 24)) @ _orth(rng, 30, 24, cplx).conj().T
    t = BlockTensor.from_blocks([a_ax, b_ax], blocks,
                                dtype=torch.complex128 if cplx else torch.float64)
    r = block_svd(t, 1, (0,), 0.0, 10 ** 6, 10 ** 6)
    spectra = {i: np.linalg.svd(b, compute_uv=False) for (i, _), b in blocks.items()}
    for (mi, _), blk in r.s.blocks().items():
        i = r.u.axes[-1].qns[mi][0] // 2
        got = np.real(np.diag(blk))
        assert np.abs(got - spectra[i][:len(got)]).max() <= 1e-13
    for (i, mi), U in r.u.blocks().items():
        assert np.abs(U.conj().T @ U - np.eye(U.shape[1])).max() < 1e-12
    for (mi, j), V in r.vt.blocks().items():
        assert np.abs(V @ V.conj().T - np.eye(V.shape[0])).max() < 1e-12
    rec = contract(contract(r.u, r.s, 1), r.vt, 1).blocks()
    for k, b in blocks.items():
        assert np.abs(rec.get(k, np.zeros_like(b)) - b).max() < 1e-10
