"""GPU-resident two-site finite DMRG (the reference's VMPS form): MPS
initialisation, environments, block SVD truncation and the sweep driver.

Reference behaviour followed:
* ``TwoSiteFiniteVMPS`` / ``TwoSiteFiniteVMPSSweep`` / ``TwoSiteFiniteVMPSUpdate``
  (``include/qlmps/algorithm/vmps/two_site_update_finite_vmps_impl.h:95-373``):
  right-moving updates i = 0..N-3, left-moving i = N-1..2; psi0 = A_l . A_r;
  Lanczos; SVD with the left divergence of A_l; the environment next to the
  updated pair is refreshed; one log line per update.
* trivial boundary environments (``vmps_init.h:157-187``).
* ``DirectStateInitMps`` (``one_dim_tn/mps/finite_mps/finite_mps_init.h:213-262``).

Everything stays on the GPU: contractions and environment updates go through
the library's planners, the Krylov solve through ``qlg_lanczos_gs``, and the
per-sector SVDs through batched FP64 eigen-decompositions of the Gram matrix
on each sector matrix's smaller side (torch / rocSOLVER).  The truncation
rule is this package's: pooled singular values, smallest D whose discarded
weight sum(s^2)/total <= trunc_err, clamped to [Dmin, Dmax].
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .mpo import Model, physical_axis
from .solver import EffectiveHamiltonian, LanczosParams
from . import _lib as L
from .tensor import IN, OUT, Axis, BlockTensor, _flow, contract, default_context, env_update


def divergence(t: BlockTensor) -> Tuple[int, ...]:
    return _flow(t.axes, t.coords[0])


def block_view(t: BlockTensor, i: int) -> torch.Tensor:
    c = t.coords[i]
    return t.data[t.offsets[i]:t.offsets[i] + t.block_size(c)].view(t.block_shape(c))


# ---------------------------------------------------------------------------
# block SVD with truncation
# ---------------------------------------------------------------------------
@dataclass
class SVDResult:
    u: BlockTensor
    s: BlockTensor
    vt: BlockTensor
    trunc_err: float
    D: int
    entropy: float


# padded sizes of the batched eigen-decompositions.  The divide-and-conquer
# solver's tridiagonal reduction is bound by host launch latency: a batch
# costs ~22 us x its padded order x (1 + 0.045 x batch) on MI355X.  Batches run
# concurrently on four lanes, but lanes contend for the launch path, so a few
# coarse buckets beat fine ones (30 ms against 38 ms per D=2000 SVD for 64-wide
# steps above 512).  Sectors with one row need no solver.
_BUCKETS = (1, 256, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192)
_SV_NOISE = 1e-6   # singular values at or below this fraction of their sector's largest count as zero


def _bucket(n: int) -> int:
    for b in _BUCKETS:
        if n <= b:
            return b
    return -(-n // 4096) * 4096


class _SvdStructure:
    """Sector grouping of a tensor's block structure for a block SVD, and the
    flat gather index that places every element of the tensor into padded,
    per-size-bucket batches of sector matrices.  Each sector matrix is stored
    with its shorter side first (transposed when it has more rows than
    columns), so its Gram matrix on that side is one batched product.  The
    host builds a per-block table; the per-element index is computed on the
    device from it."""

    def __init__(self, t: BlockTensor, ldims: int, ldiv):
        left, right = t.axes[:ldims], t.axes[ldims:]
        nb, qlen = len(t.coords), len(ldiv)
        cs = np.asarray(t.coords, dtype=np.int64).reshape(nb, t.rank)
        # per block: left charge (the sector), row / column extents, and the
        # left / right coordinates as mixed-radix keys (their integer order is
        # the tuples' lexicographic order)
        q = np.tile(np.asarray(ldiv, dtype=np.int64), (nb, 1))
        nr, nc = np.ones(nb, np.int64), np.ones(nb, np.int64)
        lkey, rkey = np.zeros(nb, np.int64), np.zeros(nb, np.int64)
        lrad = rrad = 1
        for k, ax in enumerate(t.axes):
            col = cs[:, k]
            deg = np.asarray(ax.degs, dtype=np.int64)[col]
            if k < ldims:
                sg = -1 if ax.dir == OUT else 1
                q += sg * np.asarray(ax.qns, dtype=np.int64).reshape(ax.nsect, qlen)[col]
                nr *= deg
                lkey, lrad = lkey * ax.nsect + col, lrad * ax.nsect
            else:
                nc *= deg
                rkey, rrad = rkey * ax.nsect + col, rrad * ax.nsect
        qs, gid = np.unique(q, axis=0, return_inverse=True)
        gid = gid.reshape(-1)
        ng = len(qs)

        def offsets(key, rad, ext):
            # distinct (sector, key) pairs in sorted order; each pair's offset
            # within its sector (exclusive prefix sum of extents) and each
            # sector's total extent
            pk, first, inv = np.unique(gid * rad + key, return_index=True, return_inverse=True)
            pg, pext = pk // rad, ext[first]
            ex = np.cumsum(pext) - pext
            start = np.searchsorted(pg, np.arange(ng))
            return first, pg, pext, ex - ex[start][pg], np.add.reduceat(pext, start), inv.reshape(-1)

        rfirst, rg, rext, roff, Rs, rinv = offsets(lkey, lrad, nr)
        cfirst, cg, cext, coff, Cs, cinv = offsets(rkey, rrad, nc)
        self.qs = [tuple(int(v) for v in x) for x in qs]
        self.sect = [dict(q=qq, rows={}, cols={}, rks=[], cks=[], ro={}, co={}, R=int(Rs[g]), C=int(Cs[g]),
                          tr=bool(Rs[g] > Cs[g]), n=int(min(Rs[g], Cs[g])), m=int(max(Rs[g], Cs[g])))
                     for g, qq in enumerate(self.qs)]
        for i, g, e, o in zip(rfirst.tolist(), rg.tolist(), rext.tolist(), roff.tolist()):
            sc, lk = self.sect[g], t.coords[i][:ldims]
            sc["rks"].append(lk)
            sc["rows"][lk], sc["ro"][lk] = e, o
        for i, g, e, o in zip(cfirst.tolist(), cg.tolist(), cext.tolist(), coff.tolist()):
            sc, rk = self.sect[g], t.coords[i][ldims:]
            sc["cks"].append(rk)
            sc["cols"][rk], sc["co"][rk] = e, o
        # buckets of sectors by padded short side; slot j of bucket b.  A
        # bucket is padded only to its largest member (a multiple of 16), not
        # to the bucket's bound: the eigensolver's time grows with the order.
        raw: Dict[int, List[int]] = {}
        for si, sc in enumerate(self.sect):
            raw.setdefault(_bucket(sc["n"]), []).append(si)
        self.buckets: Dict[int, List[int]] = {}
        for b, sis in raw.items():
            if b > 1:
                b = -(-max(self.sect[si]["n"] for si in sis) // 16) * 16
            self.buckets[b] = sis
        self.mpad = {b: max(16, -(-max(self.sect[si]["m"] for si in sis) // 16) * 16)
                     for b, sis in self.buckets.items()}
        # per-block table: data offset, rows, cols, batch base, row/col origin in
        # the sector matrix, transposed flag, padded long side
        sbase = np.zeros(ng, np.int64)
        smp = np.zeros(ng, np.int64)
        base = 0
        self.bucket_base = {}
        for b, sis in self.buckets.items():
            mp = self.mpad[b]
            self.bucket_base[b] = base
            for j, si in enumerate(sis):
                sbase[si] = base + j * b * mp
                smp[si] = mp
            base += len(sis) * b * mp
        tr = np.array([sc["tr"] for sc in self.sect], dtype=np.int64)
        tab = np.stack([np.asarray(t.offsets, dtype=np.int64).reshape(nb), nr, nc, sbase[gid], roff[rinv],
                        coff[cinv], tr[gid], smp[gid]], axis=1)
        self.total = base
        dev = t.data.device
        sizes = tab[:, 1] * tab[:, 2]
        ntot = int(sizes.sum())
        T = torch.from_numpy(np.ascontiguousarray(tab.T)).to(dev)           # [8, nblocks]
        size = T[1] * T[2]
        blk = torch.repeat_interleave(torch.arange(nb, device=dev), size, output_size=ntot)
        Tb = T[:, blk]                                                        # [8, elements]
        loc = torch.arange(ntot, device=dev) - (torch.cumsum(size, 0) - size)[blk]
        r = torch.div(loc, Tb[2], rounding_mode="floor")
        rr, cc = Tb[4] + r, Tb[5] + (loc - r * Tb[2])
        pos = Tb[3] + torch.where(Tb[6] > 0, cc * Tb[7] + rr, rr * Tb[7] + cc)
        dst = torch.full((t.data.numel(),), base, dtype=torch.int64, device=dev)   # base: one trash slot
        dst[Tb[0] + loc] = pos
        self.dst = dst


_SVD_CACHE: Dict[tuple, _SvdStructure] = {}


def _svd_structure(t: BlockTensor, ldims: int, ldiv) -> _SvdStructure:
    key = (tuple(t.axes), tuple(t.coords), tuple(t.offsets), ldims, tuple(ldiv), t.data.numel())
    st = _SVD_CACHE.get(key)
    if st is None:
        if len(_SVD_CACHE) > 256:
            _SVD_CACHE.clear()
        st = _SVD_CACHE[key] = _SvdStructure(t, ldims, ldiv)
    return st


_REFLECTORS: Dict[tuple, torch.Tensor] = {}


def _reflectors(b: int, dtype, device, attempt: int = 0) -> torch.Tensor:
    """Two fixed random unit vectors u_1, u_2 of length b ([2, b, 1]); the
    Householder reflectors H_i = I - 2 u_i u_i^H are Hermitian and unitary, and
    Q = H_2 H_1 turns a block-diagonal Hermitian matrix into a dense one."""
    key = (b, dtype, str(device), attempt)
    u = _REFLECTORS.get(key)
    if u is None:
        if len(_REFLECTORS) > 512:
            _REFLECTORS.clear()
        gen = torch.Generator(device="cpu").manual_seed(4321 + 7919 * attempt + b)
        z = torch.complex(torch.randn(2, b, 1, generator=gen, dtype=torch.float64),
                          torch.randn(2, b, 1, generator=gen, dtype=torch.float64))
        z = z / torch.linalg.vector_norm(z, dim=1, keepdim=True)
        u = _REFLECTORS[key] = z.to(device=device, dtype=dtype)
    return u


def _rotate_hermitian(G: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """Q G Q^H for Q = H_2 H_1 and a batch G [B, n, n] of Hermitian matrices:
    H G H = G - 2 u w^H - 2 w u^H + 4 (u^H w) u u^H with w = G u."""
    for i in range(u.shape[0]):
        ui = u[i]
        w = G @ ui                                   # [B, n, 1]
        sc = (ui.mH @ w).real                        # [B, 1, 1]
        G = G - 2 * ui @ w.mH - 2 * w @ ui.mH + 4 * sc * (ui @ ui.mH)
    return G


def _unrotate_vectors(V: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """Q^H V = H_1 H_2 V: eigenvectors of Q G Q^H back to those of G."""
    for i in reversed(range(u.shape[0])):
        ui = u[i]
        V = V - 2 * ui @ (ui.mH @ V)
    return V


def eigh_batched(mats: List[torch.Tensor]) -> List[Tuple[torch.Tensor, torch.Tensor]]:
    """Eigen-decompositions of batches of Hermitian matrices ([B, n, n] each, all
    of one dtype, n > 1) through ``qlg_eigh_batched``: one library call for all
    batches.  Returns (w [B, n] ascending, V [B, n, n] with eigenvectors as
    columns).  The inputs are consumed (they hold the eigenvectors on return)."""
    if not mats:
        return []
    dev, dt = mats[0].device, mats[0].dtype
    ctx = default_context(dev.index or 0)
    groups = (L.EighGroup * len(mats))()
    ws = []
    for g, A in zip(groups, mats):
        assert A.is_contiguous() and A.dim() == 3 and A.shape[1] == A.shape[2] and A.dtype == dt
        w = torch.empty(A.shape[0], A.shape[1], dtype=torch.float64, device=dev)
        g.n, g.batch, g.a, g.w = A.shape[1], A.shape[0], A.data_ptr(), w.data_ptr()
        ws.append(w)
    torch.cuda.synchronize()     # the library works on its own streams; it returns after the work is done
    ctx.check(ctx.lib.qlg_eigh_batched(ctx.h, L.QLG_C128 if dt.is_complex else L.QLG_F64, len(mats), groups))
    # row j of each buffer is eigenvector j (conjugated for complex data)
    return [(w, A.mT.conj()) for w, A in zip(ws, mats)]


def _eigh_grams(grams: Dict[int, torch.Tensor]) -> Dict[int, Tuple[torch.Tensor, torch.Tensor]]:
    """Eigen-decompositions of the padded Gram batches of a block SVD.

    Complex batches are rotated by a fixed unitary Q = H_2 H_1 (two random
    Householder reflectors), Q G Q^H, before the solve and the vectors rotated
    back: the complex divide-and-conquer solver returns NaN vectors when the
    tridiagonal reduction meets exactly-zero sub-columns, which block-diagonal
    (padded or reducible) Gram matrices produce.  A batch that still comes back
    with NaN is solved again with other reflectors; a third failure raises."""
    out, todo = {}, []
    for b, G in grams.items():
        if b == 1:
            out[b] = (G[:, :, 0].real.contiguous(), torch.ones_like(G))
        else:
            todo.append(b)
    if not todo:
        return out
    if not grams[todo[0]].is_complex():
        for b, r in zip(todo, eigh_batched([grams[b].clone() for b in todo])):
            out[b] = r
        return out
    for attempt in range(3):
        us = {b: _reflectors(b, grams[b].dtype, grams[b].device, attempt) for b in todo}
        res = eigh_batched([_rotate_hermitian(grams[b], us[b]).contiguous() for b in todo])
        left = []
        for b, (w, V) in zip(todo, res):
            if bool(torch.isnan(V).any() or torch.isnan(w).any()):
                left.append(b)
            else:
                out[b] = (w, _unrotate_vectors(V, us[b]))
        todo = left
        if not todo:
            return out
    raise RuntimeError(f"block SVD: the eigensolver returned NaN for Gram batches of order {todo}")


def _complete_rows(A: torch.Tensor, bad: torch.Tensor) -> torch.Tensor:
    """Replace the rows of A flagged in ``bad`` by unit vectors orthogonal to the others."""
    k, n = A.shape
    good = A[~bad]
    gen = torch.Generator(device="cpu").manual_seed(1234)
    R = torch.randn(int(bad.sum()), n, generator=gen, dtype=torch.float64).to(A.device, A.dtype)
    Q, _ = torch.linalg.qr(torch.cat([good, R]).mT)
    out = A.clone()
    out[bad] = Q.mT[good.shape[0]:]
    return out


def _gather_runs(src: torch.Tensor, starts, sizes, dev) -> torch.Tensor:
    """Concatenation of the runs src[starts[i] : starts[i] + sizes[i]] (one device gather)."""
    n = torch.tensor(sizes, dtype=torch.int64)
    total = int(n.sum())
    if total == 0:
        return torch.zeros(1, dtype=src.dtype, device=dev)
    shift = (torch.tensor(starts, dtype=torch.int64) - (torch.cumsum(n, 0) - n)).to(dev)
    blk = torch.repeat_interleave(torch.arange(len(sizes), device=dev), n.to(dev), output_size=total)
    return src[shift[blk] + torch.arange(total, device=dev)]


def _gather_cols(src: torch.Tensor, starts, rows, cols, ld, dev) -> torch.Tensor:
    """Concatenation of the row-major [rows[i] x cols[i]] sub-matrices starting at
    src[starts[i]] with leading dimension ld[i] (one device gather)."""
    size = [r * c for r, c in zip(rows, cols)]
    n = torch.tensor(size, dtype=torch.int64)
    total = int(n.sum())
    if total == 0:
        return torch.zeros(1, dtype=src.dtype, device=dev)
    T = torch.tensor([starts, cols, ld, list(np.cumsum(size) - np.array(size))], dtype=torch.int64).to(dev)
    blk = torch.repeat_interleave(torch.arange(len(size), device=dev), n.to(dev), output_size=total)
    loc = torch.arange(total, device=dev) - T[3][blk]
    c = T[1][blk]
    r = torch.div(loc, c, rounding_mode="floor")
    return src[T[0][blk] + r * T[2][blk] + (loc - r * c)]


def _gather_strided(src: torch.Tensor, tab: np.ndarray, dev) -> torch.Tensor:
    """Concatenation of the [rows x cols] matrices (row-major) whose element
    (r, c) is src[base + r * rs + c * cs], one per row (rows, cols, base, rs,
    cs) of ``tab`` (one device gather)."""
    size = tab[:, 0] * tab[:, 1]
    total = int(size.sum())
    if total == 0:
        return torch.zeros(1, dtype=src.dtype, device=dev)
    T = torch.from_numpy(np.ascontiguousarray(np.concatenate([tab, (np.cumsum(size) - size)[:, None]], axis=1).T)).to(dev)
    blk = torch.repeat_interleave(torch.arange(len(tab), device=dev), T[0] * T[1], output_size=total)
    loc = torch.arange(total, device=dev) - T[5][blk]
    c = T[1][blk]
    r = torch.div(loc, c, rounding_mode="floor")
    return src[T[2][blk] + r * T[3][blk] + (loc - r * c) * T[4][blk]]


def block_svd(t: BlockTensor, ldims: int, ldiv, trunc_err: float, Dmin: int, Dmax: int) -> SVDResult:
    """Block SVD with truncation, through FP64 eigen-decompositions of each
    sector matrix's Gram matrix on its shorter side, batched by size bucket.

    Truncation (this package's rule): pool the singular values of all sectors,
    keep the smallest D whose discarded weight sum(s^2)/total <= trunc_err,
    clamped to [Dmin, Dmax].  The Gram route resolves singular values to an
    absolute ~sqrt(n eps) of the sector's largest, so values at or below 1e-6
    of it (Schmidt weight <= 1e-12 relative) are returned as zero, and the
    singular vectors of kept zero values are completed to an orthonormal set."""
    left, right = t.axes[:ldims], t.axes[ldims:]
    dev, dt = t.data.device, t.data.dtype
    st = _svd_structure(t, ldims, ldiv)
    flat = torch.zeros(st.total + 1, dtype=dt, device=dev)
    flat[st.dst] = t.data
    per = {}   # bucket -> (E [B, b, b] descending eigenvectors, S [B, b], P = E^T X / S [B, b, mpad])
    svals = []
    Xs, grams, scales = {}, {}, {}
    for b, sis in st.buckets.items():
        mp, B = st.mpad[b], len(sis)
        X = flat[st.bucket_base[b]:st.bucket_base[b] + B * b * mp].view(B, b, mp)
        # each sector matrix is scaled to unit Frobenius norm: the solver's
        # tolerances are relative to the (padded) matrix norm, and the padding
        # below must sit at the sector's own scale, not at that of the whole state
        nrm = torch.linalg.vector_norm(X, dim=(1, 2))
        nrm = torch.where(nrm > 0, nrm, torch.ones_like(nrm))
        X = X / nrm[:, None, None].to(dt)
        scales[b] = nrm
        G = X @ X.mT.conj()
        ns = torch.tensor([st.sect[si]["n"] for si in sis], device=dev)
        pad = torch.arange(b, device=dev)[None, :] >= ns[:, None]
        # padded rows separate below the spectrum (distinct values for complex
        # data, whose solver is given a rotated, dense matrix)
        padv = -1.0 - torch.arange(b, device=dev, dtype=torch.float64) / b if G.is_complex() else -1.0
        G.diagonal(dim1=1, dim2=2)[pad] = (padv.expand(len(sis), b)[pad].to(G.dtype) if G.is_complex() else padv)
        Xs[b], grams[b] = X, G
    eig = _eigh_grams(grams)
    for b in st.buckets:
        X = Xs[b]
        w, V = eig[b]
        E = V.flip(-1)                                   # eigenvectors, descending eigenvalue
        R = E.mT.conj() @ X                              # row j = s_j v_j^H (un-normalised)
        # Singular values from the back-projected rows, s_j = |X^H e_j|: relative
        # error ~ eps s_max / s_j, against eps (s_max / s_j)^2 for sqrt(lambda_j).
        # Re-sorted descending within each sector (near-degenerate small values
        # may swap).  Values at or below _SV_NOISE of the sector's largest count
        # as zero; their singular vectors are completed to an orthonormal set below.
        S = torch.linalg.vector_norm(R, dim=-1)
        # padded positions (the last b - n eigenvectors) stay last
        nsec = torch.tensor([st.sect[si]["n"] for si in st.buckets[b]], device=dev)
        S = torch.where(torch.arange(S.shape[1], device=dev)[None, :] < nsec[:, None], S, torch.full_like(S, -1.0))
        S, order = torch.sort(S, dim=-1, descending=True)
        S = S.clamp_min(0.0)
        E = torch.gather(E, 2, order[:, None, :].expand(E.shape))
        R = torch.gather(R, 1, order[:, :, None].expand(R.shape))
        S = torch.where(S > _SV_NOISE * S[:, :1], S, torch.zeros_like(S))
        Sinv = torch.where(S > 0, 1.0 / torch.where(S > 0, S, torch.ones_like(S)), torch.zeros_like(S))
        P = R * Sinv[..., None].to(dt)                   # unit rows; zero rows for zero values
        S = S * scales[b][:, None]                       # back to the tensor's scale
        per[b] = (E, S, P)
        svals.append(S)
    # Orthogonality between rows degrades as eps s_max^2 / (s_i s_j), ~1e-4 at
    # the noise threshold: Newton-Schulz steps toward the polar factor,
    # P <- 1.5 P - 0.5 (P P^H) P, restore it quadratically.  The reconstruction
    # moves by at most ~eps s_max^2 / s.  Zero rows stay zero.  The buckets run
    # in lock step, one host check per round for all of them; each stops at
    # its own deviation < 1e-13.
    masks = {b: (per[b][1] > 0).to(dt) for b in st.buckets}
    active = list(st.buckets)
    for _ in range(8):
        if not active:
            break
        PPs = {b: per[b][2] @ per[b][2].mH for b in active}
        devs = torch.stack([(PPs[b] - torch.diag_embed(masks[b])).abs().amax() for b in active]).tolist()
        still = []
        for b, odev in zip(active, devs):
            if odev < 1e-13:
                continue
            if odev > 0.5:
                raise RuntimeError(f"block SVD: singular vectors too far from orthonormal ({odev:.2e})")
            E, S, P = per[b]
            per[b] = (E, S, 1.5 * P - 0.5 * (PPs[b] @ P))
            still.append(b)
        active = still
    # one host transfer for the truncation decision
    flat_sv = torch.cat([s_.reshape(-1) for s_ in svals]).double().cpu().numpy()
    sv_all, pos = [], 0
    for s_ in svals:
        sv_all.append(flat_sv[pos:pos + s_.numel()].reshape(s_.shape))
        pos += s_.numel()
    sv, owner, sv_host = [], [], {}
    for (b, sis), arr in zip(st.buckets.items(), sv_all):
        for j, si in enumerate(sis):
            n = st.sect[si]["n"]
            sv_host[si] = arr[j, :n]
            sv.append(arr[j, :n])
            owner.append(np.full(n, si))
    sv = np.concatenate(sv)
    owner = np.concatenate(owner)
    order = np.argsort(-sv, kind="stable")
    sq = sv[order] ** 2
    total = float(sq.sum())
    tail = np.concatenate([np.cumsum(sq[::-1])[::-1], [0.0]])
    D = len(sv)
    for k in range(len(sv) + 1):
        if tail[k] / total <= trunc_err:
            D = k
            break
    D = max(min(max(D, min(Dmin, len(sv))), Dmax, len(sv)), 1)
    err = float(tail[D] / total)
    keep = np.bincount(owner[order[:D]], minlength=len(st.sect))
    smax = float(sv[order[0]]) if len(sv) else 0.0
    kept_si = [si for si in range(len(st.sect)) if keep[si] > 0]
    mid_out = Axis.make([(st.sect[si]["q"], int(keep[si])) for si in kept_si], OUT)
    mid_in = mid_out.inverse()
    slot = {}
    for b, sis in st.buckets.items():
        for j, si in enumerate(sis):
            slot[si] = (b, j)
    # Per kept sector, U (R x k) and V^H (k x C), row-major, are gathered from
    # the buckets' E and P into two flat buffers (one device gather each), and
    # its k singular values into a third; the blocks of u and vt are then cut
    # out of them by one device gather each (u blocks are row ranges of U, vt
    # blocks column ranges of V^H), and s is scattered onto its diagonal blocks.
    #   X = M (R <= C):   U = E[:n, :k],        V^H = P[:k, :m]
    #   X = M^T (R > C):  E = conj(V), so V^H = E[:n, :k]^T and U = M conj(E) / S = P[:k, :m]^T
    parts, sparts, ebase, pbase, sbase_ = [], [], {}, {}, {}
    off = soff = 0
    for b in st.buckets:
        E, S, P = per[b]
        ebase[b], pbase[b], sbase_[b] = off, off + E.numel(), soff
        parts += [E.reshape(-1), P.reshape(-1)]
        sparts.append(S.reshape(-1))
        off += E.numel() + P.numel()
        soff += S.numel()
    utab = np.zeros((len(kept_si), 5), dtype=np.int64)
    vtab = np.zeros((len(kept_si), 5), dtype=np.int64)
    stab = np.zeros((len(kept_si), 5), dtype=np.int64)
    ubase, vbase, ks = {}, {}, {}
    uoff = voff = 0
    for mi, si in enumerate(kept_si):
        sc, k = st.sect[si], int(keep[si])
        b, j = slot[si]
        n, m, mp = sc["n"], sc["m"], st.mpad[b]
        eb, pb = ebase[b] + j * b * b, pbase[b] + j * b * mp
        if sc["tr"]:
            utab[mi], vtab[mi] = (m, k, pb, 1, mp), (k, n, eb, 1, b)
        else:
            utab[mi], vtab[mi] = (n, k, eb, b, 1), (k, m, pb, mp, 1)
        stab[mi] = (1, k, sbase_[b] + j * b, 0, 1)
        ubase[si], vbase[si], ks[si] = uoff, voff, k
        uoff += sc["R"] * k
        voff += k * sc["C"]
    src = torch.cat(parts)
    Uflat = _gather_strided(src, utab, dev).to(dt)
    Vflat = _gather_strided(src, vtab, dev).to(dt)
    Sflat = _gather_strided(torch.cat(sparts), stab, dev)
    for si in kept_si:
        k = ks[si]
        if sv_host[si][k - 1] != 0.0:
            continue
        # kept zero singular values: complete their singular vectors
        sc = st.sect[si]
        b, j = slot[si]
        bad = per[b][1][j, :k] == 0
        if sc["tr"]:
            U = Uflat[ubase[si]:ubase[si] + sc["R"] * k].view(sc["R"], k)
            U.copy_(_complete_rows(U.mT.contiguous(), bad).mT)
        else:
            Vh = Vflat[vbase[si]:vbase[si] + k * sc["C"]].view(k, sc["C"])
            Vh.copy_(_complete_rows(Vh.contiguous(), bad))
    # u blocks (lk, mi): rows [ro, ro + rows) of U -> one contiguous run each
    ublk = sorted((lk + (mi,), si, lk) for mi, si in enumerate(kept_si) for lk in st.sect[si]["rks"])
    u_start = [ubase[si] + st.sect[si]["ro"][lk] * ks[si] for _, si, lk in ublk]
    u_size = [st.sect[si]["rows"][lk] * ks[si] for _, si, lk in ublk]
    u = BlockTensor(list(left) + [mid_out], [c for c, _, _ in ublk],
                    _gather_runs(Uflat, u_start, u_size, dev))
    # vt blocks (mi, rk): columns [co, co + cols) of V^H (k x C)
    vblk = sorted(((mi,) + rk, si, rk) for mi, si in enumerate(kept_si) for rk in st.sect[si]["cks"])
    vt = BlockTensor([mid_in] + list(right), [c for c, _, _ in vblk],
                     _gather_cols(Vflat, [vbase[si] + st.sect[si]["co"][rk] for _, si, rk in vblk],
                                  [ks[si] for _, si, _ in vblk], [st.sect[si]["cols"][rk] for _, si, rk in vblk],
                                  [st.sect[si]["C"] for _, si, _ in vblk], dev))
    # s: diagonal k x k blocks in sector order
    kk = [ks[si] for si in kept_si]
    sdata = torch.zeros(max(sum(k * k for k in kk), 1), dtype=dt, device=dev)
    sb = np.concatenate([[0], np.cumsum([k * k for k in kk])[:-1]]).astype(np.int64)
    diag_pos = torch.from_numpy(np.concatenate([sb[i] + np.arange(k) * (k + 1) for i, k in enumerate(kk)])).to(dev)
    sdata[diag_pos] = Sflat.to(dt)
    s = BlockTensor([mid_in, mid_out], [(mi, mi) for mi in range(len(kept_si))], sdata)
    pr = np.sort(sv)[::-1][:D] ** 2
    pr = pr / pr.sum()
    pr = pr[pr > 0]
    ee = float(-(pr * np.log(pr)).sum())
    return SVDResult(u, s, vt, err, D, ee)


# ---------------------------------------------------------------------------
# MPS construction
# ---------------------------------------------------------------------------
def direct_state_mps(m: Model, stat_labs: Sequence[int], device="cuda", dtype=torch.float64) -> List[BlockTensor]:
    """``finite_mps_init.h:213-262``: product state, site 0 carries the total charge."""
    n = m.n
    qat = m.qns_at
    zero = tuple(0 for _ in qat(0)[0])
    total = zero
    for i in range(n):
        total = tuple(a + b for a, b in zip(total, qat(i)[stat_labs[i]]))
    mps = []
    lvb = Axis.make([(zero, 1)], IN)
    lq = zero
    for i in range(n):
        pout = physical_axis(qat(i), OUT)
        if i == 0:
            rq = tuple(a - b for a, b in zip(total, qat(0)[stat_labs[0]]))
        elif i < n - 1:
            rq = tuple(a - b for a, b in zip(lq, qat(i)[stat_labs[i]]))
        else:
            rq = zero
        rvb = Axis.make([(rq, 1)], OUT)
        blk = np.ones((1, 1, 1))
        mps.append(BlockTensor.from_blocks([lvb, pout, rvb], {(0, stat_labs[i], 0): blk}, device=device,
                                           dtype=dtype))
        lvb = rvb.inverse()
        lq = rq
    return mps


def _charge_counts(qns, nsites):
    cnt = {tuple(0 for _ in qns[0]): 1}
    for _ in range(nsites):
        new = {}
        for q, c in cnt.items():
            for s in qns:
                k = tuple(a + b for a, b in zip(q, s))
                new[k] = new.get(k, 0) + c
        cnt = new
    return cnt


def random_bonds(m: Model, total, D: int) -> List[Axis]:
    """Virtual bond sectors of a random MPS in sector ``total``: every charge a
    bond can carry, with degeneracy min(count from the left, count from the
    right), scaled down to <= D in total."""
    qns, n = m.qns, m.n
    total = tuple(total)
    bonds = []
    for b in range(n - 1):
        cl, cr = _charge_counts(qns, b + 1), _charge_counts(qns, n - b - 1)
        degs = {}
        for x, c in cl.items():
            q = tuple(a - y for a, y in zip(total, x))
            if q in cr:
                degs[q] = min(c, cr[q])
        tot = sum(degs.values())
        if tot > D:
            degs = {q: max(1, int(round(v * D / tot))) for q, v in degs.items()}
            while sum(degs.values()) > D:
                qm = max(degs, key=lambda q: degs[q])
                degs[qm] -= 1
        bonds.append(Axis.make(sorted(degs.items()), OUT))
    return bonds


def _site_axes(m: Model, total, bonds: List[Axis]):
    """(axes, divergence) of every site tensor for the given bonds."""
    pout = physical_axis(m.qns, OUT)
    zero = tuple(0 for _ in m.qns[0])
    lvb = Axis.make([(zero, 1)], IN)
    out = []
    for i in range(m.n):
        rvb = bonds[i] if i < m.n - 1 else Axis.make([(zero, 1)], OUT)
        out.append(([lvb, pout, rvb], tuple(total) if i == 0 else zero))
        lvb = rvb.inverse()
    return out


def random_mps(m: Model, total, D: int, seed: int, device="cuda", dtype=torch.float64,
               canonical: bool = True) -> List[BlockTensor]:
    """Seeded random MPS in sector ``total`` (bond dimension <= D), right-canonical
    from site 1, centre at site 0 (the same construction as the oracle's).
    ``canonical=False`` returns the raw site tensors, filled with U[-1, 1) by a
    seeded device generator (fast for long chains; not the oracle's stream)."""
    n = m.n
    rng = np.random.default_rng(seed)
    gen = None if canonical else torch.Generator(device=device).manual_seed(seed)
    mps = []
    for axes, div in _site_axes(m, total, random_bonds(m, total, D)):
        lvb, pout, rvb = axes
        blocks = {}
        for a in range(lvb.nsect):
            for s in range(pout.nsect):
                for c in range(rvb.nsect):
                    if _flow(axes, (a, s, c)) == div:
                        shp = (lvb.degs[a], pout.degs[s], rvb.degs[c])
                        blocks[(a, s, c)] = rng.uniform(-1, 1, size=shp) if canonical else shp
        if canonical:
            mps.append(BlockTensor.from_blocks(axes, blocks, device=device, dtype=dtype))
        else:
            keys = sorted(blocks)
            size = sum(int(np.prod(blocks[k])) for k in keys)
            data = torch.rand(max(size, 1), generator=gen, device=device, dtype=dtype) * 2 - 1
            mps.append(BlockTensor(axes, keys, data))
    if not canonical:
        return mps
    for i in range(n - 1, 0, -1):
        r = block_svd(mps[i], 1, divergence(mps[i]), 0.0, 1, 10 ** 9)
        mps[i] = r.vt
        us = contract(r.u, r.s, 1)
        mps[i - 1] = contract(mps[i - 1], us, 1)
    nrm = torch.linalg.vector_norm(mps[0].data).item()
    mps[0].data /= nrm
    return mps


def trivial_env(mps_bond: Axis, mpo_bond: Axis, device, dtype) -> BlockTensor:
    """``vmps_init.h:160-164``."""
    return BlockTensor.from_blocks([mps_bond, mpo_bond.inverse(), mps_bond.inverse()],
                                   {(0, 0, 0): np.ones((1, 1, 1))}, device=device, dtype=dtype)


# ---------------------------------------------------------------------------
# two-site VMPS
# ---------------------------------------------------------------------------
@dataclass
class SweepParams:
    """``algorithm/finite_vmps_sweep_params.h:31-49``."""
    sweeps: int
    Dmin: int
    Dmax: int
    trunc_err: float
    lancz: LanczosParams
    precision: str = "fp64"
    verbose: bool = True
    checkpoint_dir: Optional[str] = None   # MPS dumped here after every sweep (and on a stop signal)


@dataclass
class SweepStats:
    log: list = field(default_factory=list)
    sweep_seconds: list = field(default_factory=list)
    lanczos_s: float = 0.0
    svd_s: float = 0.0
    env_s: float = 0.0
    contract_s: float = 0.0
    plan_s: float = 0.0
    matvecs: int = 0
    algo_flops: float = 0.0


class TwoSiteVMPS:
    """Two-site finite DMRG with GPU-resident MPS and environments."""

    def __init__(self, mps: List[BlockTensor], mpo: List[BlockTensor], params: SweepParams):
        self.mps, self.mpo, self.p = mps, mpo, params
        n = len(mps)
        dev, dt = mps[0].data.device, mps[0].data.dtype
        self.renvs: List[Optional[BlockTensor]] = [None] * n
        self.lenvs: List[Optional[BlockTensor]] = [None] * n
        self.renvs[0] = trivial_env(mps[-1].axes[2], mpo[-1].axes[3], dev, dt)
        for i in range(1, n - 1):
            self.renvs[i] = env_update(self.renvs[i - 1], mps[n - i], mpo[n - i], "r")
        self.lenvs[0] = trivial_env(mps[0].axes[0], mpo[0].axes[0], dev, dt)
        self.stats = SweepStats()
        self.sweeps_done = 0
        self._flags = None

    def update(self, direction: str, target: int) -> float:
        n = len(self.mps)
        p, st = self.p, self.stats
        t0 = time.perf_counter()
        if direction == "r":
            l, r, llen, rlen = target, target + 1, target, n - (target + 2)
        else:
            l, r, llen, rlen = target - 1, target, target - 1, n - target - 1
        t = time.perf_counter()
        psi = contract(self.mps[l], self.mps[r], 1)
        st.contract_s += time.perf_counter() - t
        t = time.perf_counter()
        H = EffectiveHamiltonian([self.lenvs[llen], self.mpo[l], self.mpo[r], self.renvs[rlen]], psi,
                                 precision=p.precision)
        tp = time.perf_counter()
        st.plan_s += tp - t
        res = H.lanczos(psi, p.lancz)
        t_l = time.perf_counter() - tp
        st.lanczos_s += t_l
        st.matvecs += res.iters
        st.algo_flops += res.iters * H.flops()[0]
        H.close()
        t = time.perf_counter()
        sv = block_svd(res.gs_vec, 2, divergence(self.mps[l]), p.trunc_err, p.Dmin, p.Dmax)
        st.svd_s += time.perf_counter() - t
        t = time.perf_counter()
        if direction == "r":
            self.mps[l] = sv.u
            self.mps[r] = contract(sv.s, sv.vt, 1)
        else:
            self.mps[l] = contract(sv.u, sv.s, 1)
            self.mps[r] = sv.vt
        st.contract_s += time.perf_counter() - t
        t = time.perf_counter()
        if direction == "r":
            self.lenvs[llen + 1] = env_update(self.lenvs[llen], self.mps[target], self.mpo[target], "l")
        else:
            self.renvs[rlen + 1] = env_update(self.renvs[rlen], self.mps[target], self.mpo[target], "r")
        st.env_s += time.perf_counter() - t
        tot = time.perf_counter() - t0
        st.log.append(dict(site=target, dir=direction, e0=res.gs_eng, trunc_err=sv.trunc_err, D=sv.D,
                           iters=res.iters, lanczos_s=t_l, total_s=tot, entropy=sv.entropy))
        if p.verbose:
            print(f"Site {target:4d} E0 = {res.gs_eng:20.16f} TruncErr = {sv.trunc_err:.2e} D = {sv.D:5d} "
                  f"Iter = {res.iters:3d} LanczT = {t_l:8.4f} TotT = {tot:8.4f} S = {sv.entropy:10.7f}",
                  file=sys.stderr, flush=True)
        return res.gs_eng

    def sweep(self) -> Optional[float]:
        """One right and one left half-sweep; returns the last update's energy, or
        None after an emergency stop (the MPS is dumped mid-sweep)."""
        n = len(self.mps)
        e0 = None
        t0 = time.perf_counter()
        for direction, sites in (("r", range(0, n - 2)), ("l", range(n - 1, 1, -1))):
            for i in sites:
                e0 = self.update(direction, i)
                if self._flags is not None and self._flags.emergency:
                    from .checkpoint import save_mps
                    save_mps(self.p.checkpoint_dir, self.mps, {"complete_sweep": False,
                                                               "sweeps_done": self.sweeps_done,
                                                               "stopped_at": [direction, i], "energy": e0})
                    return None
        torch.cuda.synchronize()
        self.stats.sweep_seconds.append(time.perf_counter() - t0)
        return e0

    def run(self) -> float:
        """Sweeps until ``params.sweeps`` are done in total (a resumed driver
        continues from its ``sweeps_done``).  With ``checkpoint_dir`` set, the MPS is dumped
        after every sweep; SIGUSR1 / SIGUSR2 stop after the current sweep and
        SIGINT / SIGTERM after the current update (``dmrg_impl.h:478-502``), in
        both cases with the state dumped first."""
        from .checkpoint import StopFlags, save_mps
        e0 = None
        flags = StopFlags().install() if self.p.checkpoint_dir else None
        self._flags = flags
        try:
            for k in range(self.sweeps_done, self.p.sweeps):
                e0 = self.sweep()
                if e0 is None:        # emergency stop inside the sweep (already dumped)
                    break
                self.sweeps_done += 1
                if self.p.checkpoint_dir:
                    save_mps(self.p.checkpoint_dir, self.mps,
                             {"complete_sweep": True, "sweeps_done": self.sweeps_done, "energy": e0})
                if flags and flags.stop:
                    break
        finally:
            if flags:
                flags.uninstall()
            self._flags = None
        return e0

    @classmethod
    def resume(cls, checkpoint_dir: str, mpo: List[BlockTensor], params: SweepParams, device="cuda"):
        """Continue from a dump.  A dump taken inside a sweep is right-canonicalised
        (centre moved to site 0) first; the state itself is unchanged."""
        from .checkpoint import load_mps
        mps, man = load_mps(checkpoint_dir, device)
        if not man.get("complete_sweep", False):
            center_to_zero(mps)
        drv = cls(mps, mpo, params)
        drv.sweeps_done = int(man.get("sweeps_done", 0))
        return drv


def center_to_zero(mps: List[BlockTensor], start: Optional[int] = None):
    """Right-canonicalise sites start .. 1 (default N-1 .. 1) by SVD without
    truncation: every singular value is kept (Dmin = the full rank), so bonds
    keep their dimension and the state is unchanged."""
    n = len(mps)
    for i in range(n - 1 if start is None else start, 0, -1):
        r = block_svd(mps[i], 1, divergence(mps[i]), 0.0, 10 ** 9, 10 ** 9)
        mps[i] = r.vt
        mps[i - 1] = contract(mps[i - 1], contract(r.u, r.s, 1), 1)


def TwoSiteFiniteVMPS(mps, mpo, params: SweepParams):
    """Run ``params.sweeps`` sweeps; returns (E0, driver)."""
    drv = TwoSiteVMPS(mps, mpo, params)
    return drv.run(), drv
