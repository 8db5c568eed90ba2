"""The Krylov solvers' run-ahead form: the tridiagonal work of every iteration
and the stop rule on the GPU (FP64), host looks at the stop flag once per chunk
of queued iterations, optional hipGraph capture of the chunks, optional full
reorthogonalisation.

Bars: the run-ahead solve and the iteration-by-iteration host form
(QLG_KRYLOV_HOST=1) stop at the same iteration and return the same energy to
1e-12 and the same vector to 1e-10 in overlap, for every chunk size; graph
replays are bit-identical to the eager run-ahead solve; expmv matches the
dense exp(-i tau H) psi to 1e-10 and the oracle's iteration count; with
reorthogonalisation the energy matches dense diagonalisation (the reference
KAT, 1e-8) and a split-BF16 solve lands within 1e-9 of FP64."""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle.bsten import BSTensor, random_fill
from oracle.lanczos import LanczosExpmvSolver as OExpmv, LanczosParams as OParams, eff_ham_mul_two_site_state
from oracle.models import heisenberg_chain, hubbard_cylinder
from tests import refkat
from tests.convert import to_device, to_oracle
from tests.test_gpu_parity import _dense_two_site, _real_env_problem

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.solver import EffectiveHamiltonian, LanczosParams


class _env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update({k: str(v) for k, v in self.kv.items()})

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _overlap(a, b):
    va, vb = a.data, b.data
    return float(torch.dot(va, vb) / (torch.linalg.vector_norm(va) * torch.linalg.vector_norm(vb)))


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
def test_run_ahead_equals_host_iteration(which):
    print(f"[case] run-ahead vs host lanczos {which}", flush=True)
    if which == "heisenberg":
        eff, psi = _real_env_problem(heisenberg_chain(14), 14, 24, 1001, (0,))
    else:
        eff, psi = _real_env_problem(hubbard_cylinder(3, 2), 6, 24, 1002, (6, 0))
    d = [to_device(t) for t in eff]
    x = to_device(psi)
    H = EffectiveHamiltonian(d, x)
    prm = LanczosParams(1e-10, 100)
    with _env(QLG_KRYLOV_HOST=1):
        ref = H.lanczos(x, prm)
    for chunk in (1, 3, 8, 64):
        with _env(QLG_LANCZOS_CHUNK=chunk):
            got = H.lanczos(x, prm)
        print(f"  chunk {chunk}: iters {got.iters} (host {ref.iters}) E {got.gs_eng:.15f}", flush=True)
        assert got.iters == ref.iters
        assert abs(got.gs_eng - ref.gs_eng) <= 1e-12 * abs(ref.gs_eng)
        assert abs(abs(_overlap(got.gs_vec, ref.gs_vec)) - 1.0) < 1e-10
    H.close()


@pytest.mark.parametrize("params", [(1e-9, 200), (1e-16, 20)])
def test_run_ahead_reference_kat(params):
    """test_lanczos_solver.cc:114-167: |E - dsyev| < 1e-8 for both stop rules,
    with and without graph capture, and the host form's iteration count."""
    print(f"[case] run-ahead KAT {params}", flush=True)
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2)
    psi = random_fill(BSTensor([Din, dout, dout, Dout]), (0,), np.random.default_rng(0))
    d = [to_device(t) for t in (L, Wl, Wr, R)]
    x = to_device(psi)
    H = EffectiveHamiltonian(d, x)
    w0 = np.linalg.eigvalsh(_dense_two_site(L, Wl, Wr, R))[0]
    with _env(QLG_KRYLOV_HOST=1):
        host = H.lanczos(x, LanczosParams(*params))
    eager = H.lanczos(x, LanczosParams(*params))
    H.use_graphs(True)
    replays = []
    for _ in range(3):
        replays.append(H.lanczos(x, LanczosParams(*params)))
        torch.cuda.synchronize()
        _ = H @ x          # eager work between replays
    for r in [eager] + replays:
        assert abs(r.gs_eng - w0) < 1e-8
        assert r.iters == host.iters
    for r in replays:
        assert r.gs_eng == eager.gs_eng and torch.equal(r.gs_vec.data, eager.gs_vec.data)
    H.close()


def _complex_problem():
    refkat.srand(0)
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(12, 2, 2, complex_=True)
    psi = random_fill(BSTensor([Din, dout, dout, Dout], np.complex128), (0,), np.random.default_rng(3),
                      complex_=True)
    return [L, Wl, Wr, R], psi


@pytest.mark.parametrize("graphs", [False, True])
def test_run_ahead_expmv(graphs):
    """exp(-i tau H) psi on the device-side exponential: dense expm to 1e-10
    (the Krylov tolerance is 1e-12), the oracle's iteration count, the host
    form to 1e-13; replays identical."""
    print(f"[case] run-ahead expmv graphs={graphs}", flush=True)
    eff, psi = _complex_problem()
    Hd = _dense_two_site(*eff)
    tau = 0.3
    exact = sla.expm(-1j * tau * Hd) @ psi.to_dense().reshape(-1)
    o = OExpmv(eff, psi, eff_ham_mul_two_site_state, tau, OParams(1e-12, 30))
    d = [to_device(t) for t in eff]
    x = to_device(psi)
    H = EffectiveHamiltonian(d, x)
    with _env(QLG_KRYLOV_HOST=1):
        host = H.expmv(x, tau, LanczosParams(1e-12, 30))
    if graphs:
        H.use_graphs(True)
    runs = [H.expmv(x, tau, LanczosParams(1e-12, 30)) for _ in range(3 if graphs else 1)]
    for r in runs:
        got = to_oracle(r.expmv, psi).to_dense().reshape(-1)
        err = float(np.linalg.norm(got - exact) / np.linalg.norm(exact))
        diff = (r.expmv.data - host.expmv.data).abs().max().item() / host.expmv.data.abs().max().item()
        print(f"  iters {r.iters} (oracle {o.iters}, host {host.iters}) err vs expm {err:.2e} vs host {diff:.2e}",
              flush=True)
        assert err < 1e-10          # the Krylov stop at Distance < 1e-12
        assert r.iters == o.iters == host.iters
        assert diff <= 1e-13
    for r in runs[1:]:
        assert torch.equal(r.expmv.data, runs[0].expmv.data)
    H.close()


@pytest.mark.parametrize("passes", [1, 2])
def test_reorthogonalisation(passes):
    """Full reorthogonalisation (classical Gram-Schmidt against the whole basis):
    the reference KAT energy to 1e-8 of dense dsyev; on real environments a
    split-BF16 (3-term) solve with reorthogonalisation lands within 1e-9 of
    the FP64 energy; expmv with reorthogonalisation matches dense expm."""
    print(f"[case] reorth passes={passes}", flush=True)
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2)
    psi = random_fill(BSTensor([Din, dout, dout, Dout]), (0,), np.random.default_rng(0))
    d = [to_device(t) for t in (L, Wl, Wr, R)]
    x = to_device(psi)
    H = EffectiveHamiltonian(d, x)
    H.set_reorth(passes)
    r = H.lanczos(x, LanczosParams(1e-12, 200))
    w0 = np.linalg.eigvalsh(_dense_two_site(L, Wl, Wr, R))[0]
    assert abs(r.gs_eng - w0) < 1e-8
    H.close()
    eff, psi = _real_env_problem(hubbard_cylinder(3, 2), 6, 24, 1002, (6, 0))
    d = [to_device(t) for t in eff]
    x = to_device(psi)
    e64 = EffectiveHamiltonian(d, x).lanczos(x, LanczosParams(1e-12, 100)).gs_eng
    Hs = EffectiveHamiltonian(d, x, precision="bf16x3")
    Hs.set_reorth(passes)
    rs = Hs.lanczos(x, LanczosParams(1e-12, 100))
    print(f"  E64 {e64:.15f} bf16x3+reorth {rs.gs_eng:.15f} iters {rs.iters}", flush=True)
    assert abs(rs.gs_eng - e64) <= 1e-9 * abs(e64)
    Hs.close()
    eff, psi = _complex_problem()
    Hd = _dense_two_site(*eff)
    exact = sla.expm(-1j * 0.3 * Hd) @ psi.to_dense().reshape(-1)
    Hc = EffectiveHamiltonian([to_device(t) for t in eff], to_device(psi))
    Hc.set_reorth(passes)
    rc = Hc.expmv(to_device(psi), 0.3, LanczosParams(1e-12, 30))
    got = to_oracle(rc.expmv, psi).to_dense().reshape(-1)
    assert float(np.linalg.norm(got - exact) / np.linalg.norm(exact)) < 1e-10
    Hc.close()
