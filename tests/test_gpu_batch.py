"""Batching form: independent random-field Heisenberg chains solved together,
checked against the oracle at the configuration's own shapes.

Workload: 32 chains, L ~ U[32, 128], J = 1, fields h_i ~ U(-2, 2), D drawn from
{128, 256, 512, 768}, seed 1004 (``ultradmrg_amd.chains``); the centre two-site
effective Hamiltonian of each.  The batch plan runs every matvec stage of all
chains as one grouped launch; the batched Lanczos advances all chains in lock
step (hipGraph capture of its iterations optional).

Bars, per chain: the batched matvec equals the oracle's four-step contraction
on that chain's tensors within 1e-12 (max-abs error over max-abs value); the
batched Lanczos energy equals the oracle's LanczosSolver on the same inputs
within 1e-10 relative with the same iteration count; a graph-replayed second
solve reproduces the first bit for bit."""
import numpy as np
import pytest
import torch

from oracle.lanczos import LanczosParams as OParams, LanczosSolver as OLanczos, eff_ham_mul_two_site_state
from tests.convert import from_device

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.chains import random_field_chains
    from ultradmrg_amd.solver import BatchedEffectiveHamiltonian, EffectiveHamiltonian, LanczosParams

NCHAINS, GROUPS = 32, 4
PARAMS = (1e-10, 40)
_STATE = {}


def _batch():
    """Chains, their oracle tensors, the batched matvec output and two batched
    (graph-captured, then replayed) Lanczos solves, computed once."""
    if not _STATE:
        chains = random_field_chains(NCHAINS, seed=1004)
        hams = [EffectiveHamiltonian(c.eff_ham, c.psi) for c in chains]
        batch = BatchedEffectiveHamiltonian(hams, graphs=True)
        x = batch.pack([c.psi for c in chains])
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        batch.matvec_into(x, y)
        prm = LanczosParams(*PARAMS)
        runs = [batch.lanczos([c.psi for c in chains], prm) for _ in range(2)]
        _STATE.update(chains=chains, hams=hams, batch=batch, y=y.cpu(), runs=runs)
    return _STATE


def test_config_shapes():
    s = _batch()
    ch = s["chains"]
    assert len(ch) == NCHAINS
    assert all(32 <= c.length <= 128 for c in ch)
    assert {c.D for c in ch} <= {128, 256, 512, 768} and len({c.D for c in ch}) >= 3
    a, _ = s["batch"].flops()
    assert abs(a - sum(h.flops()[0] for h in s["hams"])) <= 1e-9 * a


@pytest.mark.parametrize("group", range(GROUPS))
def test_batched_matvec_vs_oracle(group):
    s = _batch()
    for i in range(group, NCHAINS, GROUPS):
        c, h, o = s["chains"][i], s["hams"][i], s["batch"].offsets[i]
        print(f"[case] batch matvec chain {i} L={c.length} D={c.D}", flush=True)
        oeff, opsi = [from_device(t) for t in c.eff_ham], from_device(c.psi)
        ref = eff_ham_mul_two_site_state(oeff, opsi)
        assert h.layout_coords == sorted(ref.blocks)
        got = s["y"][o:o + h.size].numpy()
        refv = np.concatenate([ref.blocks[k].reshape(-1) for k in h.layout_coords])
        assert np.abs(got - refv).max() <= 1e-12 * np.abs(refv).max()


@pytest.mark.parametrize("group", range(GROUPS))
def test_batched_lanczos_vs_oracle(group):
    s = _batch()
    first, second = s["runs"]
    for i in range(group, NCHAINS, GROUPS):
        c = s["chains"][i]
        print(f"[case] batch lanczos chain {i} L={c.length} D={c.D}", flush=True)
        oeff, opsi = [from_device(t) for t in c.eff_ham], from_device(c.psi)
        ores = OLanczos(oeff, opsi, eff_ham_mul_two_site_state, OParams(*PARAMS))
        b = first[i]
        assert b.iters == ores.iters
        assert abs(b.gs_eng - ores.gs_eng) <= 1e-10 * abs(ores.gs_eng)
        assert second[i].iters == b.iters and second[i].gs_eng == b.gs_eng
        assert torch.equal(second[i].gs_vec.data, b.gs_vec.data)


def test_batched_lanczos_equals_separate_solves():
    s = _batch()
    prm = LanczosParams(*PARAMS)
    for i in range(0, NCHAINS, 3):
        print(f"[case] batch vs separate solve chain {i}", flush=True)
        h, c, b = s["hams"][i], s["chains"][i], s["runs"][0][i]
        single = h.lanczos(c.psi, prm)
        assert b.iters == single.iters
        assert abs(b.gs_eng - single.gs_eng) <= 1e-12 * abs(single.gs_eng)
        va, vb = single.gs_vec.data[:b.gs_vec.data.numel()], b.gs_vec.data
        ov = float(torch.dot(va, vb)) / float(torch.linalg.vector_norm(va) * torch.linalg.vector_norm(vb))
        assert abs(abs(ov) - 1.0) < 1e-10
