"""Seeded block-sparse effective-Hamiltonian workloads shared by tests and the bench.

Bond sectors follow the charge bookkeeping of a finite MPS (site 0 carries
the total charge, boundary bonds are one-dimensional with zero charge); the
environment blocks are filled with U[-1, 1) so that every structurally
allowed block is present.
"""
import numpy as np

from oracle.bsten import OUT, BSTensor, Index
from oracle.models import build_mpo, heisenberg_chain, hubbard_cylinder, phys_index


def gaussian_bond(qns_site, width, D, nsect_half, qn_len=1):
    """Sector profile exp(-q^2/(2 width^2)) over charges -K..K (U(1)), scaled to total D."""
    if qn_len == 1:
        qs = [(2 * k,) for k in range(-nsect_half, nsect_half + 1)]
        w = np.array([np.exp(-(q[0] / 2) ** 2 / (2 * width ** 2)) for q in qs])
    else:
        # (N, 2Sz) with N + 2Sz even: one-electron moves change both by one unit
        qs = [(n, m) for n in range(-nsect_half, nsect_half + 1) for m in range(-nsect_half, nsect_half + 1)
              if (n + m) % 2 == 0]
        w = np.array([np.exp(-(q[0] ** 2 + q[1] ** 2) / (2 * width ** 2)) for q in qs])
    degs = np.maximum(1, np.round(w / w.sum() * D)).astype(int)
    return [(q, int(d)) for q, d in zip(qs, degs)]


def synthetic_eff_ham(kind, D, seed, nsect_half=None, width=None, site_model=None):
    """Random L, Wl, Wr, R, psi at bond dimension ~D.

    kind = "heisenberg": U(1) spin chain MPO (w=5); kind = "hubbard": U(1)xU(1)
    4-leg cylinder MPO taken from the middle of a 4x16 snake.
    """
    rng = np.random.default_rng(seed)
    if kind == "heisenberg":
        m = heisenberg_chain(8)
        mpo = build_mpo(m)
        Wl, Wr = mpo[3], mpo[4]
        nsect_half = nsect_half or 12
        width = width or max(1.0, nsect_half / 2.5)
        secs = gaussian_bond(m.qns, width, D, nsect_half, 1)
    else:
        m = hubbard_cylinder(16, 4) if site_model is None else site_model
        mpo = build_mpo(m)
        mid = m.n // 2
        Wl, Wr = mpo[mid - 1], mpo[mid]
        nsect_half = nsect_half or 5
        width = width or 2.0
        secs = gaussian_bond(m.qns, width, D, nsect_half, 2)
    zero = tuple(0 for _ in m.qns[0])
    a_out = Index.make(secs, OUT)        # right bond of the left block
    a_in = a_out.inverse()
    b_out = Index.make(secs, OUT)
    b_in = b_out.inverse()
    pout = phys_index(m.qns, OUT)
    L = BSTensor([a_in, Wl.indexes[0].inverse(), a_out], np.float64, div=zero)
    R = BSTensor([b_out, Wr.indexes[3].inverse(), b_in], np.float64, div=zero)
    for t in (L, R):
        for c in t.allowed_blocks(zero):
            t.blocks[c] = rng.uniform(-1, 1, size=t.blk_shape(c))
    psi = BSTensor([a_in, pout, pout, b_out], np.float64, div=zero)
    for c in psi.allowed_blocks(zero):
        psi.blocks[c] = rng.uniform(-1, 1, size=psi.blk_shape(c))
    return [L, Wl, Wr, R], psi


# t-J model of tests/test_algorithm/test_finite_dmrg.cc:446-583 (tJ_operators.h:110-128).
# Site basis (up, down, empty) with U(1)xU(1) charges (N, 2Sz); operators stored op[s_in, s_out].
TJ_QNS = [(1, 1), (1, -1), (0, 0)]


def tj_operators():
    def mk(entries):
        o = np.zeros((3, 3))
        for (i, j), v in entries.items():
            o[i, j] = v
        return o
    return {"bupc": mk({(2, 0): 1}), "bupa": mk({(0, 2): 1}), "bdnc": mk({(2, 1): 1}), "bdna": mk({(1, 2): 1}),
            "f": mk({(0, 0): -1, (1, 1): -1, (2, 2): 1}), "sz": mk({(0, 0): 0.5, (1, 1): -0.5}),
            "sp": mk({(1, 0): 1}), "sm": mk({(0, 1): 1}), "id": np.eye(3)}


def tj_model(model_cls, n, bonds, t, J):
    """Hopping -t (c+_i c_j + h.c.) with the Jordan-Wigner string between the sites (no double
    occupancy, so b+ F = b+ on the first site) and the Heisenberg coupling J S_i.S_j on each bond."""
    m = model_cls(n, TJ_QNS, tj_operators())
    for a, b in bonds:
        a, b = min(a, b), max(a, b)
        for o1, o2 in (("bupc", "bupa"), ("bdnc", "bdna"), ("bupa", "bupc"), ("bdna", "bdnc")):
            m.add(-t, (a, o1), (b, o2), string="f")
        m.add(J, (a, "sz"), (b, "sz"))
        m.add(J / 2, (a, "sp"), (b, "sm"))
        m.add(J / 2, (a, "sm"), (b, "sp"))
    return m


TJ_1D_BONDS = [(0, 1), (1, 2), (2, 3)]
TJ_2D_BONDS = [(0, 1), (0, 2), (2, 3), (1, 3)]


# Rashba t-J model of tests/test_algorithm/test_finite_dmrg.cc:598-706: charge-only U(1) (N), site basis
# (up, down, empty), t=3, J=1, spin-orbit lambda=0.03 (imaginary along y), 3x2 lattice periodic in x.
RASHBA_QNS = [(1,), (1,), (0,)]


def rashba_terms(t=3.0, J=1.0, lamb=0.03, Nx=3, Ny=2):
    """(coef, op_a, site_a, op_b, site_b, string) terms of test_finite_dmrg.cc:620-706."""
    il = 1j * lamb
    idx = lambda x, y: x * Ny + y
    T = []
    for x in range(Nx):
        for y in range(Ny):
            for kind in ("x", "y"):
                if kind == "x":
                    s0, s1 = sorted((idx(x, y), idx((x + 1) % Nx, y)))      # BCx periodic
                    so = ([(lamb, "cdagup", "cdn"), (lamb, "cup", "cdagdn"), (-lamb, "cdagdn", "cup"), (-lamb, "cdn", "cdagup")]
                          if x != Nx - 1 else
                          [(lamb, "cdn", "cdagup"), (lamb, "cdagdn", "cup"), (-lamb, "cup", "cdagdn"), (-lamb, "cdagup", "cdn")])
                else:
                    if y == Ny - 1:                                          # BCy open
                        continue
                    s0, s1 = sorted((idx(x, y), idx(x, (y + 1) % Ny)))
                    so = ([(il, "cdagup", "cdn"), (il, "cdagdn", "cup"), (-il, "cup", "cdagdn"), (-il, "cdn", "cdagup")]
                          if y != Ny - 1 else
                          [(il, "cup", "cdagdn"), (il, "cdn", "cdagup"), (-il, "cdagup", "cdn"), (-il, "cdagdn", "cup")])
                for c, oa, ob in [(-t, "cdagup", "cup"), (-t, "cdagdn", "cdn"), (-t, "cup", "cdagup"), (-t, "cdn", "cdagdn")] + so:
                    T.append((c, oa, s0, ob, s1, "f"))
                for c, oa, ob in [(J, "sz", "sz"), (J / 2, "sp", "sm"), (J / 2, "sm", "sp"), (-J / 4, "ntot", "ntot")]:
                    T.append((c, oa, s0, ob, s1, None))
    return T

def _std(entries):
    """Matrix <out|O|in> over (up, down, empty)."""
    o = np.zeros((3, 3), complex)
    for (out, inn), v in entries.items():
        o[out, inn] = v
    return o


RASHBA_STD_OPS = {"cup": _std({(2, 0): 1}), "cdagup": _std({(0, 2): 1}), "cdn": _std({(2, 1): 1}),
                  "cdagdn": _std({(1, 2): 1}), "f": _std({(0, 0): -1, (1, 1): -1, (2, 2): 1}),
                  "sz": _std({(0, 0): .5, (1, 1): -.5}), "sp": _std({(0, 1): 1}), "sm": _std({(1, 0): 1}),
                  "ntot": _std({(0, 0): 1, (1, 1): 1})}


# Kondo double chain of tests/test_algorithm/test_finite_dmrg.cc:846-952: sites of two kinds, an
# extended electron (basis up, down, double, empty) on even sites and a localized spin (up, down) on
# odd sites; no conserved charge is used (every state has charge 0).  Operators as <out|O|in>.
def _mat(d, entries):
    o = np.zeros((d, d))
    for (out, inn), v in entries.items():
        o[out, inn] = v
    return o


KONDO_STD_OPS = {
    "sz": _mat(4, {(0, 0): 0.5, (1, 1): -0.5}), "sp": _mat(4, {(0, 1): 1}), "sm": _mat(4, {(1, 0): 1}),
    "bupcF": _mat(4, {(2, 1): -1, (0, 3): 1}), "bupaF": _mat(4, {(1, 2): 1, (3, 0): -1}),
    "Fbdnc": _mat(4, {(2, 0): 1, (1, 3): -1}), "Fbdna": _mat(4, {(0, 2): -1, (3, 1): 1}),
    "bupc": _mat(4, {(2, 1): 1, (0, 3): 1}), "bupa": _mat(4, {(1, 2): 1, (3, 0): 1}),
    "bdnc": _mat(4, {(2, 0): 1, (1, 3): 1}), "bdna": _mat(4, {(0, 2): 1, (3, 1): 1}),
    "Sz": _mat(2, {(0, 0): 0.5, (1, 1): -0.5}), "Sp": _mat(2, {(0, 1): 1}), "Sm": _mat(2, {(1, 0): 1})}


def kondo_terms(nx=4, t=0.25, jk=1.0, jz=0.5):
    """(coef, op_a, site_a, op_b, site_b) terms of test_finite_dmrg.cc:925-937 (no string operators)."""
    n, T = 2 * nx, []
    for i in range(0, n - 2, 2):
        T += [(-t, "bupcF", i, "bupa", i + 2), (-t, "bdnc", i, "Fbdna", i + 2), (t, "bupaF", i, "bupc", i + 2),
              (t, "bdna", i, "Fbdnc", i + 2), (jz, "Sz", i + 1, "Sz", i + 3)]
    for i in range(0, n, 2):
        T += [(jk, "sz", i, "Sz", i + 1), (jk / 2, "sp", i, "Sm", i + 1), (jk / 2, "sm", i, "Sp", i + 1)]
    return T
