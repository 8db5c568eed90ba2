"""Oracle: local Hilbert spaces, operators, MPOs and exact diagonalisation for the
models the benchmark configurations use.

TEST INFRASTRUCTURE ONLY (see ``oracle/bsten.py``).

* Spin-1/2 with U(1) Sz, basis {up: 2Sz=+1, down: 2Sz=-1}
  (``include/qlmps/model_relevant/sites/spin/spin_one_half_sites.h:30-47``).
* Hubbard with U(1)xU(1) (N, 2Sz), basis {updown:(2,0), up:(1,1), down:(1,-1), empty:(0,0)}
  (``model_relevant/sites/fermion/hubbard_sites.h:30-41``); the Jordan-Wigner
  dressed operator matrices are those of ``benchmark/ultradmrg_hubbard2d.cpp:66-114``
  and ``model_relevant/operators/fermion/hubbard_operators.h:59-118``.
* Operators are stored ``op[in, out] = <out|O|in>`` (the reference's
  ``Tensor({col, row}) = <row|O|col>`` convention, ``hubbard_operators.h:17-19``).
* MPO tensors are ``W[wl(IN), s_in(IN), s_out(OUT), wr(OUT)]`` with zero
  divergence and one-dimensional zero-charge boundary bonds
  (``one_dim_tn/mpo/mpogen/mpogen_impl.h:443-504``).  The finite-state
  automaton here is this package's own: one ``start`` and one ``end`` state per
  bond plus one state per open (first site, first operator, string) prefix.
  It realises the same Hamiltonian as the reference's symbolic generator; the
  bond dimension may differ, so FLOP counts are stated per MPO.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp

from .bsten import IN, OUT, BSTensor, Index, qn_neg, qn_sub

# --------------------------------------------------------------------------
# local spaces and operators
# --------------------------------------------------------------------------
SPIN_QNS = [(1,), (-1,)]
HUBBARD_QNS = [(2, 0), (1, 1), (1, -1), (0, 0)]


def spin_ops() -> Dict[str, np.ndarray]:
    """op[in, out] for spin-1/2; index 0 = up, 1 = down."""
    z = lambda: np.zeros((2, 2))
    sz, spl, smi, idm = z(), z(), z(), z()
    sz[0, 0], sz[1, 1] = 0.5, -0.5
    spl[1, 0] = 1.0          # <up|S+|down>
    smi[0, 1] = 1.0          # <down|S-|up>
    idm[0, 0] = idm[1, 1] = 1.0
    return {"sz": sz, "sp": spl, "sm": smi, "id": idm}


def hubbard_ops() -> Dict[str, np.ndarray]:
    """op[in, out] for the Hubbard site; 0=updown, 1=up, 2=down, 3=empty."""
    names = ["sz", "sp", "sm", "id", "f", "bupc", "bupa", "bdnc", "bdna", "bupcF", "bupaF",
             "Fbdnc", "Fbdna", "cupccdnc", "cdnacupa", "Uterm", "nf", "nup", "ndn"]
    o = {n: np.zeros((4, 4)) for n in names}
    o["sz"][1, 1], o["sz"][2, 2] = 0.5, -0.5
    o["sp"][2, 1] = 1.0
    o["sm"][1, 2] = 1.0
    for i in range(4):
        o["id"][i, i] = 1.0
    o["f"][0, 0], o["f"][1, 1], o["f"][2, 2], o["f"][3, 3] = 1, -1, -1, 1
    o["bupc"][2, 0] = 1; o["bupc"][3, 1] = 1
    o["bdnc"][1, 0] = 1; o["bdnc"][3, 2] = 1
    o["bupa"][0, 2] = 1; o["bupa"][1, 3] = 1
    o["bdna"][0, 1] = 1; o["bdna"][2, 3] = 1
    o["bupcF"][2, 0] = -1; o["bupcF"][3, 1] = 1
    o["Fbdnc"][1, 0] = 1; o["Fbdnc"][3, 2] = -1
    o["bupaF"][0, 2] = 1; o["bupaF"][1, 3] = -1
    o["Fbdna"][0, 1] = -1; o["Fbdna"][2, 3] = 1
    o["cupccdnc"][3, 0] = 1
    o["cdnacupa"][0, 3] = 1
    o["Uterm"][0, 0] = 1
    o["nf"][0, 0], o["nf"][1, 1], o["nf"][2, 2] = 2, 1, 1
    o["nup"][0, 0], o["nup"][1, 1] = 1, 1
    o["ndn"][0, 0], o["ndn"][2, 2] = 1, 1
    return o


def op_div(op: np.ndarray, qns) -> Tuple[int, ...]:
    """Charge carried by an operator: q_out - q_in of any nonzero element."""
    nz = np.argwhere(np.abs(op) > 0)
    if len(nz) == 0:
        return tuple(0 for _ in qns[0])
    i, o = nz[0]
    d = qn_sub(qns[o], qns[i])
    for i2, o2 in nz:
        if qn_sub(qns[o2], qns[i2]) != d:
            raise ValueError("operator does not carry a definite charge")
    return d


def phys_index(qns, dir=OUT) -> Index:
    return Index.make([(q, 1) for q in qns], dir)


# --------------------------------------------------------------------------
# terms and the finite-state-automaton MPO
# --------------------------------------------------------------------------
@dataclass
class Term:
    coef: float
    ops: List[Tuple[int, str]]           # [(site, op name)] with ascending sites (1 or 2 entries)
    string: Optional[str] = None        # operator inserted strictly between the two sites


@dataclass
class ModelSpec:
    n: int
    qns: list
    ops: Dict[str, np.ndarray]
    terms: List[Term] = field(default_factory=list)

    def add(self, coef, *site_ops, string=None):
        so = sorted(site_ops, key=lambda x: x[0])
        self.terms.append(Term(float(coef), [(int(s), o) for s, o in so], string))


def heisenberg_chain(n: int, J: float = 1.0, fields: Optional[Sequence[float]] = None) -> ModelSpec:
    """H = J sum (Sz Sz + 1/2 (S+S- + S-S+)) + sum h_i Sz_i on an open chain."""
    m = ModelSpec(n, SPIN_QNS, spin_ops())
    for i in range(n - 1):
        m.add(J, (i, "sz"), (i + 1, "sz"))
        m.add(0.5 * J, (i, "sp"), (i + 1, "sm"))
        m.add(0.5 * J, (i, "sm"), (i + 1, "sp"))
    if fields is not None:
        for i, h in enumerate(fields):
            if h != 0.0:
                m.add(h, (i, "sz"))
    return m


def _add_hopping(m: ModelSpec, t: float, i: int, j: int):
    """``hubbard_operators.h:185-197`` / ``ultradmrg_hubbard2d.cpp:157-160``."""
    s1, s2 = min(i, j), max(i, j)
    m.add(-t, (s1, "bupcF"), (s2, "bupa"), string="f")
    m.add(t, (s1, "bupaF"), (s2, "bupc"), string="f")
    m.add(-t, (s1, "bdnc"), (s2, "Fbdna"), string="f")
    m.add(t, (s1, "bdna"), (s2, "Fbdnc"), string="f")


def hubbard_cylinder(lx: int, ly: int, t: float = 1.0, U: float = 8.0, t2: float = 0.0) -> ModelSpec:
    """Snake-ordered Ly-leg cylinder, ``benchmark/ultradmrg_hubbard2d.cpp:149-202``."""
    n = lx * ly
    m = ModelSpec(n, HUBBARD_QNS, hubbard_ops())
    for i in range(n):
        m.add(U, (i, "Uterm"))
    for i in range(n - ly):
        _add_hopping(m, t, i, i + ly)
    for i in range(n):
        y = i % ly
        if y < ly - 1:
            _add_hopping(m, t, i, i + 1)
        elif ly > 2:
            _add_hopping(m, t, i - ly + 1, i)
    if t2 != 0.0:
        for i in range(n - ly):
            y, x = i % ly, i // ly
            for txy in ((y + 1) % ly + (x + 1) * ly, (y + ly - 1) % ly + (x + 1) * ly):
                _add_hopping(m, t2, i, txy)
    return m


def hubbard_square(nx: int, ny: int, t0: float, t1: float, U: float) -> ModelSpec:
    """``tests/test_algorithm/test_finite_dmrg.cc:729-760`` (site = x*Ny + y)."""
    m = ModelSpec(nx * ny, HUBBARD_QNS, hubbard_ops())
    idx = lambda x, y: x * ny + y
    for i in range(nx):
        for j in range(ny):
            s0 = idx(i, j)
            m.add(U, (s0, "Uterm"))
            if i != nx - 1:
                _add_hopping(m, t0, s0, idx(i + 1, j))
            if j != ny - 1:
                _add_hopping(m, t0, s0, idx(i, j + 1))
            if j != ny - 1:
                if i != 0:
                    _add_hopping(m, t1, s0, idx(i - 1, j + 1))
                if i != nx - 1:
                    _add_hopping(m, t1, s0, idx(i + 1, j + 1))
    return m


def build_mpo(m: ModelSpec, dtype=np.float64) -> List[BSTensor]:
    """Finite-state-automaton MPO with states grouped into charge sectors."""
    n, qns, ops = m.n, m.qns, m.ops
    d = len(qns)
    zero = tuple(0 for _ in qns[0])
    ident = np.eye(d)
    # open prefixes: key = (first site, first op, string) -> last site reached
    last_site: Dict[tuple, int] = {}
    for t in m.terms:
        if len(t.ops) == 2:
            (i, oi), (j, _) = t.ops
            key = (i, oi, t.string)
            last_site[key] = max(last_site.get(key, -1), j)
        elif len(t.ops) != 1:
            raise ValueError("only one- and two-site terms are supported")
    # bond b (right of site b): states
    bonds: List[List[tuple]] = []
    for b in range(n - 1):
        st = [("start",), ("end",)]
        for key, j in sorted(last_site.items(), key=lambda kv: (kv[0][0], kv[0][1], str(kv[0][2]))):
            if key[0] <= b < j:
                st.append(("open",) + key)
        bonds.append(st)

    def state_qn(s):
        if s[0] in ("start", "end"):
            return zero
        return qn_neg(op_div(ops[s[2]], qns))

    def bond_index(states, dir):
        groups: Dict[tuple, List[tuple]] = {}
        for s in states:
            groups.setdefault(state_qn(s), []).append(s)
        order = sorted(groups)
        secs = [(q, len(groups[q])) for q in order]
        pos = {}
        off = 0
        for q in order:
            for s in groups[q]:
                pos[s] = off
                off += 1
        return Index.make(secs, dir), pos

    phys_out = phys_index(qns, OUT)
    phys_in = phys_out.inverse()
    mpo = []
    for k in range(n):
        lstates = [("start",)] if k == 0 else bonds[k - 1]
        rstates = [("end",)] if k == n - 1 else bonds[k]
        lidx, lpos = bond_index(lstates, IN)
        ridx, rpos = bond_index(rstates, OUT)
        Wd = np.zeros((lidx.dim, d, d, ridx.dim), dtype=dtype)

        def put(ls, rs, mat):
            if ls in lpos and rs in rpos:
                Wd[lpos[ls], :, :, rpos[rs]] += mat

        put(("start",), ("start",), ident)
        put(("end",), ("end",), ident)
        for t in m.terms:
            if len(t.ops) == 1 and t.ops[0][0] == k:
                put(("start",), ("end",), t.coef * ops[t.ops[0][1]])
            elif len(t.ops) == 2:
                (i, oi), (j, oj) = t.ops
                key = ("open", i, oi, t.string)
                if i == k:
                    pass  # opening handled once per prefix below
                if j == k:
                    put(key, ("end",), t.coef * ops[oj])
        for (i, oi, s), j in last_site.items():
            key = ("open", i, oi, s)
            if i == k:
                put(("start",), key, ops[oi])
            elif i < k < j:
                put(key, key, ops[s] if s is not None else ident)
        W = BSTensor([lidx, phys_in, phys_out, ridx], dtype, div=zero)
        lo, ro = lidx.offsets(), ridx.offsets()
        for a in range(lidx.nsect):
            for si in range(d):
                for so in range(d):
                    for c in range(ridx.nsect):
                        blk = Wd[lo[a]:lo[a] + lidx.deg(a), si:si + 1, so:so + 1, ro[c]:ro[c] + ridx.deg(c)]
                        if np.any(blk != 0):
                            if W.blk_qn_flow((a, si, so, c)) != zero:
                                raise AssertionError("MPO element violates charge conservation")
                            W.blocks[(a, si, so, c)] = blk.copy()
        mpo.append(W)
    return mpo


def mpo_bond_dims(mpo: List[BSTensor]) -> List[int]:
    return [W.indexes[3].dim for W in mpo]


# --------------------------------------------------------------------------
# exact diagonalisation (small systems; pins the MPO and end-to-end energies)
# --------------------------------------------------------------------------
def dense_hamiltonian_sparse(m: ModelSpec, sector_qn: Optional[Tuple[int, ...]] = None):
    """Sparse H on the product basis (site 0 most significant), optionally
    restricted to the states with total charge ``sector_qn``."""
    n, qns, ops = m.n, m.qns, m.ops
    d = len(qns)
    # basis restriction
    if sector_qn is not None:
        labels = np.array(qns)
        tot = np.zeros((1, labels.shape[1]), dtype=np.int64)
        for _ in range(n):
            tot = (tot[:, None, :] + labels[None, :, :]).reshape(-1, labels.shape[1])
        keep = np.where(np.all(tot == np.array(sector_qn)[None, :], axis=1))[0]
    else:
        keep = None
    H = sp.csr_matrix((d ** n, d ** n))
    eye = lambda k: sp.identity(d ** k, format="csr")
    for t in m.terms:
        mats = [sp.identity(d, format="csr")] * n
        mats = list(mats)
        for s, o in t.ops:
            mats[s] = sp.csr_matrix(ops[o].T)
        if len(t.ops) == 2 and t.string is not None:
            for s in range(t.ops[0][0] + 1, t.ops[1][0]):
                mats[s] = sp.csr_matrix(ops[t.string].T)
        term = mats[0]
        for k in range(1, n):
            term = sp.kron(term, mats[k], format="csr")
        H = H + t.coef * term
    if keep is not None:
        H = H[keep][:, keep]
    return H


def ed_ground_energy(m: ModelSpec, sector_qn) -> float:
    import scipy.sparse.linalg as spla
    H = dense_hamiltonian_sparse(m, sector_qn)
    if H.shape[0] <= 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    w = spla.eigsh(H, k=1, which="SA", tol=1e-13)[0]
    return float(w[0])
