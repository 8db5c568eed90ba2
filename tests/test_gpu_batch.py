"""Batching form: independent random-field Heisenberg chains solved together.

The batch plan runs every matvec stage of all chains as one grouped launch,
and the batched Lanczos solver advances all chains in lock step.  Each chain's
result must be the one a separate solve gives: the same iteration count,
energy within 1e-12 relative and the same ground state; the batched matvec
equals each member's own matvec to 1e-13; with hipGraph capture enabled, a
second solve replays the captured iterations and reproduces the first."""
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.chains import random_field_chains
    from ultradmrg_amd.solver import BatchedEffectiveHamiltonian, EffectiveHamiltonian, LanczosParams


def _chains():
    return random_field_chains(5, seed=1004, lmin=8, lmax=16, dims=(8, 16, 24))


def test_batched_matvec_equals_members():
    print("[case] batched matvec, 5 chains", flush=True)
    chains = _chains()
    hams = [EffectiveHamiltonian(c.eff_ham, c.psi) for c in chains]
    batch = BatchedEffectiveHamiltonian(hams)
    x = batch.pack([c.psi for c in chains])
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    batch.matvec_into(x, y)
    for h, c, o in zip(hams, chains, batch.offsets):
        ref = h @ c.psi
        got = y[o:o + h.size]
        assert float((got - ref.data[:h.size]).abs().max()) <= 1e-13 * float(ref.data.abs().max())
    a, _ = batch.flops()
    assert abs(a - sum(h.flops()[0] for h in hams)) <= 1e-9 * a


@pytest.mark.parametrize("graphs", [False, True])
def test_batched_lanczos_equals_separate_solves(graphs):
    print(f"[case] batched lanczos graphs={graphs}", flush=True)
    chains = _chains()
    prm = LanczosParams(1e-10, 60)
    hams = [EffectiveHamiltonian(c.eff_ham, c.psi) for c in chains]
    single = [h.lanczos(c.psi, prm) for h, c in zip(hams, chains)]
    batch = BatchedEffectiveHamiltonian(hams, graphs=graphs)
    runs = [batch.lanczos([c.psi for c in chains], prm) for _ in range(2 if graphs else 1)]
    for res in runs:
        for s, b in zip(single, res):
            assert b.iters == s.iters
            assert abs(b.gs_eng - s.gs_eng) <= 1e-12 * abs(s.gs_eng)
            va, vb = s.gs_vec.data[:b.gs_vec.data.numel()], b.gs_vec.data
            ov = float(torch.dot(va, vb)) / float(torch.linalg.vector_norm(va) * torch.linalg.vector_norm(vb))
            assert abs(abs(ov) - 1.0) < 1e-10
    if graphs:
        for a, b in zip(runs[0], runs[1]):
            assert a.iters == b.iters and a.gs_eng == b.gs_eng
