"""Host bookkeeping of the block SVD, on the CPU.

The block SVD (sweep.block_svd; the reference's SVD of the two-site state in
two_site_update_finite_vmps_impl.h:302) gathers every charge sector's matrix,
shorter side first, into padded per-size batches, and afterwards cuts U, V^H
and the kept singular values back out of the batches.  These tests check that
bookkeeping against dense sector matrices assembled block by block, with
numpy, from the definition: a sector collects the blocks whose left legs
carry the same charge, its rows are the distinct left coordinates in sorted
order and its columns the distinct right coordinates in sorted order."""
import itertools
import random

import numpy as np
import pytest
import torch

from ultradmrg_amd.sweep import _SvdStructure, _gather_strided
from ultradmrg_amd.tensor import IN, OUT, Axis, BlockTensor


def _random_tensor(rng, rank, qlen):
    axes = []
    for _ in range(rank):
        qs = set()
        ns = rng.randint(1, 4)
        while len(qs) < ns:
            qs.add(tuple(rng.randint(-2, 2) for _ in range(qlen)))
        qs = list(qs)
        rng.shuffle(qs)
        axes.append(Axis.make([(q, rng.randint(1, 6)) for q in qs], rng.choice([IN, OUT])))
    allc = list(itertools.product(*[range(ax.nsect) for ax in axes]))
    coords = rng.sample(allc, rng.randint(1, min(len(allc), 25)))
    sizes = [int(np.prod([ax.degs[s] for ax, s in zip(axes, c)])) for c in coords]
    perm = list(range(len(coords)))
    rng.shuffle(perm)   # blocks stored in an order unrelated to their coordinates
    offsets, acc = [0] * len(coords), 0
    for i in perm:
        offsets[i], acc = acc, acc + sizes[i]
    data = torch.from_numpy(np.random.default_rng(rng.randint(0, 10**6)).standard_normal(sum(sizes)))
    return BlockTensor(axes, coords, data, offsets)


def _dense_sectors(t, ldims, ldiv):
    """charge -> dense sector matrix, from the definition."""
    blocks = t.blocks()
    groups = {}
    for c in t.coords:
        q = list(ldiv)
        for ax, s in zip(t.axes[:ldims], c[:ldims]):
            for j, v in enumerate(ax.qns[s]):
                q[j] += -v if ax.dir == OUT else v
        groups.setdefault(tuple(q), []).append(c)
    out = {}
    for q, cs in groups.items():
        lks = sorted({c[:ldims] for c in cs})
        rks = sorted({c[ldims:] for c in cs})
        rdim = {k: int(np.prod([t.axes[i].degs[s] for i, s in enumerate(k)])) for k in lks}
        cdim = {k: int(np.prod([t.axes[ldims + i].degs[s] for i, s in enumerate(k)])) for k in rks}
        ro = dict(zip(lks, np.concatenate([[0], np.cumsum([rdim[k] for k in lks])[:-1]])))
        co = dict(zip(rks, np.concatenate([[0], np.cumsum([cdim[k] for k in rks])[:-1]])))
        M = np.zeros((sum(rdim.values()), sum(cdim.values())))
        for c in cs:
            lk, rk = c[:ldims], c[ldims:]
            M[ro[lk]:ro[lk] + rdim[lk], co[rk]:co[rk] + cdim[rk]] = blocks[c].reshape(rdim[lk], cdim[rk])
        out[q] = M
    return out


@pytest.mark.parametrize("seed", range(40))
def test_sector_batches_match_dense_sectors(seed):
    rng = random.Random(seed)
    rank = rng.randint(2, 4)
    t = _random_tensor(rng, rank, rng.choice([1, 2]))
    ldims = rng.randint(1, rank - 1)
    ldiv = tuple(rng.randint(-1, 1) for _ in t.axes[0].qns[0])
    st = _SvdStructure(t, ldims, ldiv)
    dense = _dense_sectors(t, ldims, ldiv)
    assert st.qs == sorted(dense)
    flat = torch.zeros(st.total + 1, dtype=t.data.dtype)
    flat[st.dst] = t.data
    seen = 0
    for b, sis in st.buckets.items():
        mp = st.mpad[b]
        batch = flat[st.bucket_base[b]:st.bucket_base[b] + len(sis) * b * mp].view(len(sis), b, mp).numpy()
        assert b % 16 == 0 or b == 1
        for j, si in enumerate(sis):
            sc = st.sect[si]
            M = dense[sc["q"]]
            X = M.T if M.shape[0] > M.shape[1] else M     # shorter side first
            assert (sc["R"], sc["C"]) == M.shape and sc["tr"] == (M.shape[0] > M.shape[1])
            assert sc["n"] <= b and sc["m"] <= mp
            np.testing.assert_array_equal(batch[j, :X.shape[0], :X.shape[1]], X)
            assert not batch[j, X.shape[0]:, :].any() and not batch[j, :, X.shape[1]:].any()
            seen += 1
    assert seen == len(dense)


@pytest.mark.parametrize("seed", range(10))
def test_strided_gather_matches_loops(seed):
    rng = np.random.default_rng(seed)
    src = torch.from_numpy(rng.standard_normal(5000))
    rows = []
    for _ in range(rng.integers(1, 8)):
        r, c = int(rng.integers(0, 7)), int(rng.integers(1, 7))
        rs, cs = int(rng.integers(0, 40)), int(rng.integers(0, 40))
        base = int(rng.integers(0, 5000 - max(r - 1, 0) * rs - (c - 1) * cs))
        rows.append((r, c, base, rs, cs))
    tab = np.array(rows, dtype=np.int64)
    got = _gather_strided(src, tab, torch.device("cpu")).numpy()
    want = [src[base + i * rs + k * cs].item() for r, c, base, rs, cs in rows for i in range(r) for k in range(c)]
    if want:
        np.testing.assert_array_equal(got, np.array(want))
    else:
        assert got.shape == (1,) and got[0] == 0
