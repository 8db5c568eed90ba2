"""GPU two-site TDVP against the oracle's restatement of the reference's TDVP
sweep (same inputs, truncation active) and against exact evolution exp(-iHt)
(untruncated bond dimension, where two-site TDVP is exact up to the Krylov
tolerance).  Bars: local densities within 1e-9 of the oracle and of the exact
values; the doublon quench (c^dag_up c^dag_dn on the centre site, the TDVP
configuration's initial operator) on identical inputs within 1e-9."""
import numpy as np
import pytest
import torch

from oracle.bsten import BSTensor
from oracle.dmrg import init_envs, random_mps
from oracle.lanczos import LanczosParams as OParams
from oracle.models import build_mpo, dense_hamiltonian_sparse, hubbard_cylinder
from oracle import tdvp as OT
from tests.convert import to_device
from tests.test_oracle_tdvp import mps_to_vector, site_op_dense

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd import tdvp as GT
    from ultradmrg_amd.solver import LanczosParams


def _complex(mps):
    out = []
    for t in mps:
        c = BSTensor(t.indexes, np.complex128, t.div())
        c.blocks = {k: v.astype(np.complex128) for k, v in t.blocks.items()}
        out.append(c)
    nrm = np.sqrt(sum(float(np.sum(np.abs(v) ** 2)) for v in out[0].blocks.values()))
    out[0].blocks = {k: v / nrm for k, v in out[0].blocks.items()}
    return out


def test_tdvp_exact_evolution_gpu():
    print("[case] tdvp exact hubbard 2x2", flush=True)
    import scipy.sparse.linalg as spla
    m, total = hubbard_cylinder(2, 2), (4, 0)
    mps = _complex(random_mps(m, total, 16, seed=7))
    v0, keep = mps_to_vector(mps, m.qns, total)
    H = dense_hamiltonian_sparse(m, total).astype(np.complex128)
    tau, steps = 0.05, 4
    drv = GT.TwoSiteTDVP([to_device(t) for t in mps], [to_device(t) for t in build_mpo(m)],
                         GT.TDVPParams(tau, steps, 1, 10 ** 6, 0.0, LanczosParams(1e-12, 40)))
    for _ in range(steps):
        drv.sweep()
    got = drv.measure(m.ops["nf"])
    psit = spla.expm_multiply(-1j * tau * steps * H, v0[keep])
    ref = [float(np.real(np.vdot(psit, site_op_dense(m.ops["nf"], i, m.n, 4, keep) @ psit))) for i in range(m.n)]
    assert np.abs(got - np.array(ref)).max() < 1e-9, (got, ref)


def test_tdvp_doublon_quench_vs_oracle():
    print("[case] tdvp doublon quench vs oracle (2x3, D=12)", flush=True)
    m, total = hubbard_cylinder(3, 2), (4, 0)
    mpo = build_mpo(m)
    mps = _complex(random_mps(m, total, 12, seed=21))
    site = m.n // 2
    op = m.ops["cupccdnc"]
    # oracle
    o = [t.copy() for t in mps]
    OT.apply_local_op(o, _op_tensor(op, o[site].indexes[1]), site)
    nrm = np.sqrt(sum(float(np.sum(np.abs(v) ** 2)) for v in o[0].blocks.values()))
    o[0].blocks = {k: v / nrm for k, v in o[0].blocks.items()}
    lenvs, renvs = init_envs(o, mpo)
    tau, steps, Dmax = 0.05, 3, 12
    lp = OParams(1e-12, 40)
    for _ in range(steps):
        OT.tdvp_sweep(o, lenvs, renvs, mpo, tau, 1e-12, 1, Dmax, lp)
    OT.center_to_zero(o)
    ref = OT.local_expectations(o, m.ops["nf"])
    # GPU
    g = [to_device(t) for t in mps]
    GT.apply_local_op(g, op, site)
    g[0] = GT.BlockTensor(g[0].axes, g[0].coords, g[0].data / torch.linalg.vector_norm(g[0].data), g[0].offsets)
    drv = GT.TwoSiteTDVP(g, [to_device(t) for t in mpo],
                         GT.TDVPParams(tau, steps, 1, Dmax, 1e-12, LanczosParams(1e-12, 40)))
    for _ in range(steps):
        drv.sweep()
    got = drv.measure(m.ops["nf"])
    assert abs(sum(got) - 6.0) < 1e-9          # the quench adds one pair: 6 electrons
    assert np.abs(got - ref).max() < 1e-9, (got, ref)


def _op_tensor(op, phys):
    t = BSTensor([phys.inverse(), phys], np.complex128)
    offs = phys.offsets()
    for si in range(phys.nsect):
        for so in range(phys.nsect):
            blk = op[offs[si]:offs[si] + phys.deg(si), offs[so]:offs[so] + phys.deg(so)]
            if np.any(blk != 0):
                t.blocks[(si, so)] = blk.astype(np.complex128)
    return t


def test_tdvp_split_precision_vs_fp64():
    """The split-BF16 (6-term) complex form of the doublon quench (2x3, D=12,
    three steps) against the complex128 form on identical inputs: local
    densities within 1e-5."""
    print("[case] tdvp split-bf16 vs fp64", flush=True)
    m, total = hubbard_cylinder(3, 2), (4, 0)
    mpo = build_mpo(m)
    mps = _complex(random_mps(m, total, 12, seed=21))
    site = m.n // 2
    dens = {}
    for prec in ("fp64", "bf16x6"):
        g = [to_device(t) for t in mps]
        GT.apply_local_op(g, m.ops["cupccdnc"], site)
        g[0] = GT.BlockTensor(g[0].axes, g[0].coords, g[0].data / torch.linalg.vector_norm(g[0].data), g[0].offsets)
        drv = GT.TwoSiteTDVP(g, [to_device(t) for t in mpo],
                             GT.TDVPParams(0.05, 3, 1, 12, 1e-12, LanczosParams(1e-12, 40), precision=prec))
        for _ in range(3):
            drv.sweep()
        dens[prec] = drv.measure(m.ops["nf"])
    err = np.abs(dens["bf16x6"] - dens["fp64"]).max()
    print(f"  max density error {err:.2e}", flush=True)
    assert err < 1e-5


def _dense_site(t) -> np.ndarray:
    """Dense array of a device block tensor."""
    offs = [np.concatenate([[0], np.cumsum(ax.degs)]) for ax in t.axes]
    out = np.zeros([int(o[-1]) for o in offs], dtype=np.complex128)
    host = t.data.detach().cpu().numpy()
    for c, o in zip(t.coords, t.offsets):
        shp = t.block_shape(c)
        out[tuple(slice(offs[i][c[i]], offs[i][c[i]] + shp[i]) for i in range(len(c)))] = \
            host[o:o + int(np.prod(shp))].reshape(shp)
    return out


def _dense_state(mps) -> np.ndarray:
    acc = _dense_site(mps[0])
    for t in mps[1:]:
        acc = np.tensordot(acc, _dense_site(t), axes=([-1], [0]))
    return acc.reshape(-1)


def _op_string(ops_at, n, d):
    """Dense product of site operators (op[s_in, s_out], identity elsewhere); site 0 most significant."""
    import scipy.sparse as sp
    m = None
    for k in range(n):
        o = sp.csr_matrix(ops_at[k].T) if k in ops_at else sp.identity(d, format="csr")
        m = o if m is None else sp.kron(m, o, format="csr")
    return m


def _free_fermion_correlation(time, L, x1, x2):
    """test_two_site_finite_tdvp.cc:82-100: <c_x2(t) c+_x1(0)> of the open chain H = -sum c+_i c_j + h.c."""
    res = 0.0j
    for k in range(1, L + 1):
        eps = -2.0 * np.cos(k * np.pi / (L + 1))
        if eps > 0:
            res += np.exp(-1j * eps * time) * np.sin(k * np.pi / (L + 1) * (x1 + 1)) * \
                np.sin(k * np.pi / (L + 1) * (x2 + 1))
    return res * 2.0 / (L + 1)


def test_tdvp_free_fermion_dynamic_correlation():
    """test_two_site_finite_tdvp.cc:102-151: ground state of the N=6 free-fermion chain by two-site
    DMRG (complex Hamiltonian; 4 sweeps, D 1..16, LanczosParams(1e-8)), c+ applied on site N/2 with
    the Jordan-Wigner string on the sites to its left, 30 TDVP steps of tau=0.01 (D 10..16,
    truncation 1e-10, LanczosParams(1e-8)); <c_x2(t) c+_x1(0)> = e^{i E0 t} <psi0|c_x2 psi(t)> for
    every x2 at t = 0.1, 0.2, 0.3 equals the closed form to 1e-10 (real and imaginary parts).  The
    spinless chain is realised as the spin-up electrons of the U=0 Hubbard chain (N=3, 2Sz=3)."""
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, direct_state_mps
    print("[case] tdvp free-fermion dynamic correlation N=6", flush=True)
    n, x1, d = 6, 3, 4
    m = M.Model(n, M.HUBBARD_QNS, M.hubbard_operators())
    for i in range(n - 1):
        M._hopping(m, 1.0, i, i + 1)
    dt = torch.complex128
    mpo = M.build_mpo(m, dtype=dt)
    drv = TwoSiteVMPS(direct_state_mps(m, [1, 3, 1, 3, 1, 3], dtype=dt), mpo,
                      SweepParams(4, 1, 16, 1e-10, LanczosParams(1e-8), verbose=False))
    e0 = drv.run()
    assert abs(e0 - (-3.4939592074349334893668128643185)) < 1e-13
    ops = m.ops
    v0 = _dense_state(drv.mps)
    phi = list(drv.mps)
    for k in range(x1):
        GT.apply_local_op(phi, ops["f"], k)
    GT.apply_local_op(phi, ops["bupc"], x1)
    cdag = _op_string({**{k: ops["f"] for k in range(x1)}, x1: ops["bupc"]}, n, d)
    nrm0 = np.linalg.norm(cdag @ v0)
    assert abs(np.linalg.norm(_dense_state(phi)) - nrm0) < 1e-12
    c_at = [_op_string({**{k: ops["f"] for k in range(x2)}, x2: ops["bupa"]}, n, d) for x2 in range(n)]
    td = GT.TwoSiteTDVP(phi, mpo, GT.TDVPParams(0.01, 30, 10, 16, 1e-10, LanczosParams(1e-8)))
    worst = 0.0
    for step in range(1, 31):
        td.sweep()
        if step % 10:
            continue
        t = 0.01 * step
        vt = _dense_state(td.mps)
        assert abs(np.linalg.norm(vt) - nrm0) < 1e-9
        for x2 in range(n):
            g = np.exp(1j * e0 * t) * np.vdot(v0, c_at[x2] @ vt)
            ref = _free_fermion_correlation(t, n, x1, x2)
            worst = max(worst, abs(g.real - ref.real), abs(g.imag - ref.imag))
    print(f"[free fermion] worst deviation from the closed form {worst:.2e}", flush=True)
    assert worst < 1e-10
