"""Oracle: numpy restatement of the block-sparse symmetric tensor semantics the
reference's hot path relies on.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``ultradmrg_amd``,
``include/``) may import this module; only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only to check results.

The reference delegates all tensor arithmetic to the external, un-vendored
TensorToolkit (qlten).  What is restated here is the behaviour the reference's
call sites depend on:

* ``Index``: ordered list of quantum-number sectors ``(qn, degeneracy)`` with a
  direction IN/OUT; the direction flips under ``InverseIndex``/``Dag``
  (call sites: ``include/qlmps/algorithm/vmps/vmps_init.h:160-163``,
  ``tests/test_algorithm/test_lanczos_solver.cc:34-39``).
* A block (one sector per axis) may exist only when
  ``sum_OUT qn - sum_IN qn == div`` (U(1) and U(1)xU(1) charges are additive;
  ``one_dim_tn/mps/finite_mps/finite_mps_init.h:229-258`` builds bonds with
  exactly this rule).
* ``Contract(A, B, {axesA, axesB})``: result axes are A's free axes in order
  followed by B's free axes in order (``lanczos_vmps_solver_impl.h:268-269``;
  the dense effective-Hamiltonian builder of
  ``tests/test_algorithm/test_lanczos_solver.cc:64-68`` depends on it).
* ``Contract<T, Q, a_tail, b_head>(A, B, a0, b0, n)`` (rotated form,
  ``lanczos_vmps_solver_impl.h:270-272``, ``vmps_init.h:352-371``): contracts the
  ``n`` cyclically consecutive axes of A starting at ``a0`` with those of B
  starting at ``b0``; the result holds A's free axes in cyclic order starting
  right after the contracted run, then B's free axes likewise.  The two flags
  only steer qlten's internal data movement; they do not change the result.
* A result block exists iff at least one pair of input blocks contributes to
  it (block existence is structural, not value based).
* ``LinearCombine`` forms the union of blocks; ``Dag`` conjugates data and
  inverts every direction; ``Normalize``/``QuasiNormalize`` scale by the
  inverse 2-norm and return the old norm (identical for bosonic tensors).

Parity status: pinned by the dense-matrix identity (block-sparse contraction
vs. ``numpy.einsum`` on the densified operands, tests/test_oracle_bsten.py),
by the reference's Lanczos / expmv known-answer tests, and end to end by the
reference's ground-state energies.  The output *order* of the rotated
contraction is additionally pinned by the matvec producing a vector in the
input state's own index order, which the dense check verifies.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

IN = -1
OUT = +1

QN = Tuple[int, ...]


def qn_add(a: QN, b: QN) -> QN:
    return tuple(x + y for x, y in zip(a, b))


def qn_sub(a: QN, b: QN) -> QN:
    return tuple(x - y for x, y in zip(a, b))


def qn_neg(a: QN) -> QN:
    return tuple(-x for x in a)


@dataclass(frozen=True)
class Index:
    """Ordered quantum-number sectors with a direction (IN=-1 / OUT=+1)."""
    sectors: Tuple[Tuple[QN, int], ...]
    dir: int

    @staticmethod
    def make(sectors: Sequence[Tuple[Sequence[int], int]], dir: int) -> "Index":
        return Index(tuple((tuple(int(v) for v in q), int(d)) for q, d in sectors), int(dir))

    def inverse(self) -> "Index":
        return Index(self.sectors, -self.dir)

    @property
    def nsect(self) -> int:
        return len(self.sectors)

    @property
    def dim(self) -> int:
        return sum(d for _, d in self.sectors)

    def qn(self, s: int) -> QN:
        return self.sectors[s][0]

    def deg(self, s: int) -> int:
        return self.sectors[s][1]

    def offsets(self) -> List[int]:
        out, acc = [], 0
        for _, d in self.sectors:
            out.append(acc)
            acc += d
        return out

    def sector_of_coor(self, c: int) -> Tuple[int, int]:
        """Actual coordinate -> (sector id, offset inside sector)."""
        acc = 0
        for s, (_, d) in enumerate(self.sectors):
            if c < acc + d:
                return s, c - acc
            acc += d
        raise IndexError(c)

    def qn_len(self) -> int:
        return len(self.sectors[0][0])


def inverse_index(idx: Index) -> Index:
    return idx.inverse()


class BSTensor:
    """Block-sparse tensor: ``blocks[coords] = ndarray`` with coords = sector ids."""

    def __init__(self, indexes: Sequence[Index], dtype=np.float64, div: QN | None = None):
        self.indexes: List[Index] = list(indexes)
        self.dtype = np.dtype(dtype)
        self.blocks: Dict[Tuple[int, ...], np.ndarray] = {}
        self._div = div

    # ----- structure -----------------------------------------------------
    @property
    def rank(self) -> int:
        return len(self.indexes)

    @property
    def shape(self) -> Tuple[int, ...]:
        return tuple(ix.dim for ix in self.indexes)

    def blk_qn_flow(self, coords: Sequence[int]) -> QN:
        q = None
        for ix, s in zip(self.indexes, coords):
            term = ix.qn(s) if ix.dir == OUT else qn_neg(ix.qn(s))
            q = term if q is None else qn_add(q, term)
        return q

    def div(self) -> QN | None:
        if self._div is not None:
            return self._div
        for c in self.blocks:
            return self.blk_qn_flow(c)
        return None

    def allowed_blocks(self, div: QN) -> List[Tuple[int, ...]]:
        out = []
        for coords in itertools.product(*[range(ix.nsect) for ix in self.indexes]):
            if self.blk_qn_flow(coords) == tuple(div):
                out.append(coords)
        return out

    def blk_shape(self, coords) -> Tuple[int, ...]:
        return tuple(ix.deg(s) for ix, s in zip(self.indexes, coords))

    def size(self) -> int:
        """Number of stored elements (qlten ``size()``: sum of block sizes)."""
        return int(sum(b.size for b in self.blocks.values()))

    def copy(self) -> "BSTensor":
        t = BSTensor(self.indexes, self.dtype, self._div)
        t.blocks = {k: v.copy() for k, v in self.blocks.items()}
        return t

    def sorted_keys(self):
        return sorted(self.blocks.keys())

    # ----- element access --------------------------------------------------
    def __setitem__(self, coors, value):
        coords, inner = zip(*[ix.sector_of_coor(c) for ix, c in zip(self.indexes, coors)])
        blk = self.blocks.get(coords)
        if blk is None:
            if self._div is not None and self.blk_qn_flow(coords) != tuple(self._div):
                raise ValueError("element outside the tensor's divergence sector")
            blk = np.zeros(self.blk_shape(coords), dtype=self.dtype)
            self.blocks[coords] = blk
        blk[inner] = value

    def __getitem__(self, coors):
        coords, inner = zip(*[ix.sector_of_coor(c) for ix, c in zip(self.indexes, coors)])
        blk = self.blocks.get(coords)
        return self.dtype.type(0) if blk is None else blk[inner]

    # ----- dense view (tests only) -------------------------------------------
    def to_dense(self) -> np.ndarray:
        out = np.zeros(self.shape, dtype=self.dtype)
        offs = [ix.offsets() for ix in self.indexes]
        for coords, blk in self.blocks.items():
            sl = tuple(slice(offs[i][s], offs[i][s] + self.indexes[i].deg(s)) for i, s in enumerate(coords))
            out[sl] = blk
        return out

    # ----- flat vector in canonical block order (lexicographic coords) -------
    def to_flat(self, keys=None) -> np.ndarray:
        keys = self.sorted_keys() if keys is None else keys
        parts = [self.blocks[k].reshape(-1) if k in self.blocks else
                 np.zeros(int(np.prod(self.blk_shape(k))), self.dtype) for k in keys]
        return np.concatenate(parts) if parts else np.zeros(0, self.dtype)

    def from_flat(self, flat: np.ndarray, keys) -> "BSTensor":
        t = BSTensor(self.indexes, flat.dtype, self._div)
        off = 0
        for k in keys:
            shp = t.blk_shape(k)
            n = int(np.prod(shp))
            t.blocks[k] = flat[off:off + n].reshape(shp).copy()
            off += n
        return t

    # ----- arithmetic --------------------------------------------------------
    def norm(self) -> float:
        return float(np.sqrt(sum(np.vdot(b, b).real for b in self.blocks.values())))

    def normalize(self) -> float:
        """qlten ``Normalize``/``QuasiNormalize`` (bosonic): scale to unit norm, return old norm."""
        n = self.norm()
        if n != 0.0:
            for k in self.blocks:
                self.blocks[k] = self.blocks[k] / n
        return n

    def scale(self, c):
        for k in self.blocks:
            self.blocks[k] = self.blocks[k] * c
        return self

    def transpose(self, perm: Sequence[int]) -> "BSTensor":
        t = BSTensor([self.indexes[p] for p in perm], self.dtype, self._div)
        for k, v in self.blocks.items():
            t.blocks[tuple(k[p] for p in perm)] = np.transpose(v, perm)
        return t


def Dag(t: BSTensor) -> BSTensor:
    """Conjugate the data and invert every index direction (qlten ``Dag``)."""
    d = None if t._div is None else qn_neg(t._div)
    r = BSTensor([ix.inverse() for ix in t.indexes], t.dtype, d)
    r.blocks = {k: np.conj(v) for k, v in t.blocks.items()}
    return r


def Contract(A: BSTensor, B: BSTensor, axes: Sequence[Sequence[int]]) -> BSTensor:
    """qlten ``Contract(A, B, {axesA, axesB}, C)``.

    Result axes: A's free axes (original order), then B's free axes.  Each pair
    of blocks whose contracted sectors match contributes ``tensordot``; a result
    block exists iff at least one pair feeds it.
    """
    axA, axB = list(axes[0]), list(axes[1])
    for a, b in zip(axA, axB):
        ia, ib = A.indexes[a], B.indexes[b]
        if ia.sectors != ib.sectors or ia.dir != -ib.dir:
            raise ValueError(f"cannot contract axis {a} of A with axis {b} of B")
    freeA = [i for i in range(A.rank) if i not in axA]
    freeB = [i for i in range(B.rank) if i not in axB]
    dtype = np.result_type(A.dtype, B.dtype)
    C = BSTensor([A.indexes[i] for i in freeA] + [B.indexes[i] for i in freeB], dtype)
    # bucket B blocks by contracted sectors
    bucket: Dict[Tuple[int, ...], List[Tuple[int, ...]]] = {}
    for kb in B.blocks:
        bucket.setdefault(tuple(kb[i] for i in axB), []).append(kb)
    for ka in sorted(A.blocks):
        key = tuple(ka[i] for i in axA)
        for kb in bucket.get(key, ()):
            kc = tuple(ka[i] for i in freeA) + tuple(kb[i] for i in freeB)
            prod = np.tensordot(A.blocks[ka], B.blocks[kb], axes=(axA, axB))
            if kc in C.blocks:
                C.blocks[kc] = C.blocks[kc] + prod
            else:
                C.blocks[kc] = prod.astype(dtype, copy=True)
    return C


def ContractRotated(A: BSTensor, B: BSTensor, a0: int, b0: int, n: int) -> BSTensor:
    """qlten ``Contract<TenElemT, QNT, a_tail, b_head>(A, B, a0, b0, n, C)``.

    Contracted axes are cyclic runs; the result is A's free axes in cyclic order
    starting after the run, then B's free axes likewise.
    """
    axA = [(a0 + i) % A.rank for i in range(n)]
    axB = [(b0 + i) % B.rank for i in range(n)]
    freeA = [(a0 + n + i) % A.rank for i in range(A.rank - n)]
    freeB = [(b0 + n + i) % B.rank for i in range(B.rank - n)]
    C = Contract(A, B, (axA, axB))
    # Contract produced freeA sorted ascending then freeB sorted ascending.
    natA, natB = sorted(freeA), sorted(freeB)
    perm = [natA.index(i) for i in freeA] + [len(natA) + natB.index(i) for i in freeB]
    if perm == list(range(len(perm))):
        return C
    return C.transpose(perm)


def LinearCombine(coefs: Sequence, tens: Sequence[BSTensor], beta, res: BSTensor) -> BSTensor:
    """qlten ``LinearCombine(coefs, tensors, beta, res)``: res <- beta*res + sum c_i t_i (block union)."""
    out = BSTensor(res.indexes if res.indexes else tens[0].indexes,
                   np.result_type(res.dtype, *[t.dtype for t in tens], np.asarray(coefs).dtype), res._div)
    for k, v in res.blocks.items():
        out.blocks[k] = beta * v
    for c, t in zip(coefs, tens):
        for k, v in t.blocks.items():
            if k in out.blocks:
                out.blocks[k] = out.blocks[k] + c * v
            else:
                out.blocks[k] = c * v
    res.indexes, res.blocks, res.dtype = out.indexes, out.blocks, out.dtype
    return res


def vdot(x: BSTensor, y: BSTensor):
    """``Contract(x, Dag(y), all axes)()`` = sum x * conj(y) over common blocks."""
    s = 0.0
    for k, v in x.blocks.items():
        w = y.blocks.get(k)
        if w is not None:
            s = s + np.sum(v * np.conj(w))
    return s


def random_fill(t: BSTensor, div: QN, rng: np.random.Generator, complex_: bool = False) -> BSTensor:
    """Create every block allowed by ``div`` and fill it with U[-1, 1)."""
    t._div = tuple(div)
    for coords in t.allowed_blocks(div):
        shp = t.blk_shape(coords)
        v = rng.uniform(-1.0, 1.0, size=shp)
        if complex_:
            v = v + 1j * rng.uniform(-1.0, 1.0, size=shp)
        t.blocks[coords] = v.astype(t.dtype)
    return t
