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
