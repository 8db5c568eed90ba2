"""Local Hilbert spaces, operators and finite-state-automaton MPOs for the models
the benchmark configurations run (spin-1/2 Heisenberg chains with fields,
Hubbard cylinders in snake order).

Conventions of the reference:
* spin-1/2 with U(1), basis {up: 2Sz=+1, down: 2Sz=-1}
  (``include/qlmps/model_relevant/sites/spin/spin_one_half_sites.h:30-47``);
* Hubbard with U(1)xU(1) (N, 2Sz), basis {updown, up, down, empty}
  (``model_relevant/sites/fermion/hubbard_sites.h:30-41``) and the
  Jordan-Wigner dressed operators of ``benchmark/ultradmrg_hubbard2d.cpp:66-114``;
* operator arrays are ``op[in, out] = <out|O|in>`` (``hubbard_operators.h:17-19``);
* MPO tensors ``W[wl(IN), s_in(IN), s_out(OUT), wr(OUT)]`` with zero divergence
  and one-dimensional zero-charge boundary bonds (``mpogen_impl.h:443-504``).

The automaton keeps one ``start`` and one ``end`` state per bond plus one state
per open (first site, first operator, string) prefix of a two-site term.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .tensor import IN, OUT, Axis, BlockTensor

SPIN_QNS = [(1,), (-1,)]
HUBBARD_QNS = [(2, 0), (1, 1), (1, -1), (0, 0)]


def spin_operators() -> Dict[str, np.ndarray]:
    ops = {k: np.zeros((2, 2)) for k in ("sz", "sp", "sm", "id")}
    ops["sz"][0, 0], ops["sz"][1, 1] = 0.5, -0.5
    ops["sp"][1, 0] = 1.0
    ops["sm"][0, 1] = 1.0
    ops["id"][0, 0] = ops["id"][1, 1] = 1.0
    return ops


def hubbard_operators() -> Dict[str, np.ndarray]:
    nz = {  # (in, out): value, following benchmark/ultradmrg_hubbard2d.cpp:66-114
        "sz": {(1, 1): 0.5, (2, 2): -0.5}, "sp": {(2, 1): 1.0}, "sm": {(1, 2): 1.0},
        "id": {(0, 0): 1, (1, 1): 1, (2, 2): 1, (3, 3): 1},
        "f": {(0, 0): 1, (1, 1): -1, (2, 2): -1, (3, 3): 1},
        "bupc": {(2, 0): 1, (3, 1): 1}, "bdnc": {(1, 0): 1, (3, 2): 1},
        "bupa": {(0, 2): 1, (1, 3): 1}, "bdna": {(0, 1): 1, (2, 3): 1},
        "bupcF": {(2, 0): -1, (3, 1): 1}, "Fbdnc": {(1, 0): 1, (3, 2): -1},
        "bupaF": {(0, 2): 1, (1, 3): -1}, "Fbdna": {(0, 1): -1, (2, 3): 1},
        "cupccdnc": {(3, 0): 1}, "cdnacupa": {(0, 3): 1}, "Uterm": {(0, 0): 1},
        "nf": {(0, 0): 2, (1, 1): 1, (2, 2): 1}, "nup": {(0, 0): 1, (1, 1): 1}, "ndn": {(0, 0): 1, (2, 2): 1},
    }
    out = {}
    for k, d in nz.items():
        m = np.zeros((4, 4))
        for (i, o), v in d.items():
            m[i, o] = v
        out[k] = m
    return out


@dataclass
class Term:
    coef: float           # complex for complex couplings
    ops: List[Tuple[int, str]]
    string: Optional[str] = None


@dataclass
class Model:
    n: int
    qns: list                              # charges of the site basis states
    ops: Dict[str, np.ndarray]
    terms: List[Term] = field(default_factory=list)
    site_qns: Optional[list] = None        # per-site charge lists, for sites of different kinds

    def qns_at(self, k: int) -> list:
        return self.site_qns[k] if self.site_qns is not None else self.qns

    def add(self, coef, *site_ops, string=None):
        so = sorted(site_ops, key=lambda x: x[0])
        c = complex(coef)
        self.terms.append(Term(c if c.imag != 0 else c.real, [(int(s), o) for s, o in so], string))


def heisenberg_chain(n: int, J: float = 1.0, fields: Optional[Sequence[float]] = None) -> Model:
    m = Model(n, SPIN_QNS, spin_operators())
    for i in range(n - 1):
        m.add(J, (i, "sz"), (i + 1, "sz"))
        m.add(0.5 * J, (i, "sp"), (i + 1, "sm"))
        m.add(0.5 * J, (i, "sm"), (i + 1, "sp"))
    if fields is not None:
        for i, h in enumerate(fields):
            if h != 0.0:
                m.add(h, (i, "sz"))
    return m


def _hopping(m: Model, t: float, i: int, j: int):
    """``hubbard_operators.h:185-197``."""
    a, b = min(i, j), max(i, j)
    m.add(-t, (a, "bupcF"), (b, "bupa"), string="f")
    m.add(t, (a, "bupaF"), (b, "bupc"), string="f")
    m.add(-t, (a, "bdnc"), (b, "Fbdna"), string="f")
    m.add(t, (a, "bdna"), (b, "Fbdnc"), string="f")


def hubbard_cylinder(lx: int, ly: int, t: float = 1.0, U: float = 8.0, t2: float = 0.0) -> Model:
    """``benchmark/ultradmrg_hubbard2d.cpp:149-202`` (snake order, site = x*ly + y)."""
    n = lx * ly
    m = Model(n, HUBBARD_QNS, hubbard_operators())
    for i in range(n):
        m.add(U, (i, "Uterm"))
    for i in range(n - ly):
        _hopping(m, t, i, i + ly)
    for i in range(n):
        y = i % ly
        if y < ly - 1:
            _hopping(m, t, i, i + 1)
        elif ly > 2:
            _hopping(m, t, i - ly + 1, i)
    if t2 != 0.0:
        for i in range(n - ly):
            y, x = i % ly, i // ly
            for j in ((y + 1) % ly + (x + 1) * ly, (y + ly - 1) % ly + (x + 1) * ly):
                _hopping(m, t2, i, j)
    return m


def _op_charge(op: np.ndarray, qns) -> Tuple[int, ...]:
    nz = np.argwhere(np.abs(op) > 0)
    if len(nz) == 0:
        return tuple(0 for _ in qns[0])
    i, o = nz[0]
    return tuple(a - b for a, b in zip(qns[o], qns[i]))


def physical_axis(qns, dir=OUT) -> Axis:
    return Axis.make([(q, 1) for q in qns], dir)


def build_mpo(m: Model, device="cuda", dtype=torch.float64) -> List[BlockTensor]:
    """Block-sparse MPO tensors on the device, one per site."""
    n, ops = m.n, m.ops
    zero = tuple(0 for _ in m.qns_at(0)[0])
    cplx = (any(np.iscomplexobj(o) and np.any(np.imag(o) != 0) for o in ops.values()) or
            any(isinstance(t.coef, complex) for t in m.terms))
    if cplx and not dtype.is_complex:
        raise ValueError("the model has complex operators or couplings: build its MPO with a complex dtype")
    last: Dict[tuple, int] = {}
    for t in m.terms:
        if len(t.ops) == 2:
            (i, oi), (j, _) = t.ops
            key = (i, oi, t.string)
            last[key] = max(last.get(key, -1), j)
        elif len(t.ops) != 1:
            raise ValueError("only one- and two-site terms")
    order = sorted(last.items(), key=lambda kv: (kv[0][0], kv[0][1], str(kv[0][2])))
    bonds = []
    for b in range(n - 1):
        st = [("start",), ("end",)] + [("open",) + key for key, j in order if key[0] <= b < j]
        bonds.append(st)

    def qn_of(s):
        if s[0] != "open":
            return zero
        return tuple(-v for v in _op_charge(ops[s[2]], m.qns_at(s[1])))

    def bond_axis(states, dir):
        groups: Dict[tuple, list] = {}
        for s in states:
            groups.setdefault(qn_of(s), []).append(s)
        pos, off, secs = {}, 0, []
        for q in sorted(groups):
            secs.append((q, len(groups[q])))
            for s in groups[q]:
                pos[s] = off
                off += 1
        return Axis.make(secs, dir), pos

    out = []
    for k in range(n):
        qns = m.qns_at(k)
        d = len(qns)
        pout = physical_axis(qns, OUT)
        pin = pout.inverse()
        ls = [("start",)] if k == 0 else bonds[k - 1]
        rs = [("end",)] if k == n - 1 else bonds[k]
        la, lp = bond_axis(ls, IN)
        ra, rp = bond_axis(rs, OUT)
        Wd = np.zeros((la.dim, d, d, ra.dim), dtype=np.complex128 if cplx else np.float64)

        def put(a, b, mat):
            if a in lp and b in rp:
                Wd[lp[a], :, :, rp[b]] += mat

        put(("start",), ("start",), np.eye(d))
        put(("end",), ("end",), np.eye(d))
        for t in m.terms:
            if len(t.ops) == 1 and t.ops[0][0] == k:
                put(("start",), ("end",), t.coef * ops[t.ops[0][1]])
            elif len(t.ops) == 2 and t.ops[1][0] == k:
                (i, oi), (_, oj) = t.ops
                put(("open", i, oi, t.string), ("end",), t.coef * ops[oj])
        for (i, oi, s), j in last.items():
            key = ("open", i, oi, s)
            if i == k:
                put(("start",), key, ops[oi])
            elif i < k < j:
                put(key, key, ops[s] if s is not None else np.eye(d))
        blocks = {}
        lo = np.concatenate([[0], np.cumsum(la.degs)])
        ro = np.concatenate([[0], np.cumsum(ra.degs)])
        for a in range(la.nsect):
            for si in range(d):
                for so in range(d):
                    for c in range(ra.nsect):
                        blk = Wd[lo[a]:lo[a + 1], si:si + 1, so:so + 1, ro[c]:ro[c + 1]]
                        if np.any(blk != 0):
                            blocks[(a, si, so, c)] = blk
        out.append(BlockTensor.from_blocks([la, pin, pout, ra], blocks, device=device, dtype=dtype))
    return out
