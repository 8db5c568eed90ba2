"""The header-only C++ layer (include/qlmps_amd.h) driven by a C++ program that
replays the reference's Lanczos known-answer test with the reference's own call
LanczosSolver(eff_ham, init, &eff_ham_mul_two_site_state, params)."""
import os
import subprocess

import numpy as np
import pytest

from oracle.bsten import BSTensor
from oracle.lanczos import LanczosParams as OParams, LanczosSolver as OLanczos, TridiagGsSolver, \
    eff_ham_mul_two_site_state
from tests import refkat

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "header_kat")


def test_header_layer_lanczos_kat():
    print("[case] C++ header KAT", flush=True)
    assert os.path.exists(EXE), "build the C++ test programs first (python -m ultradmrg_amd.build)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.splitlines()]
    runs = [(float(l[1]), int(l[3]), float(l[5])) for l in lines if l[0] == "E"]
    assert len(runs) == 2
    sb = [(float(l[1]), int(l[3])) for l in lines if l[0] == "SB"]
    assert len(sb) == 1
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(20, 2, 2)
    t = np.einsum("xwa,wsty->xatsy", L.to_dense(), Wl.to_dense())
    t = np.einsum("xatsy,yuvz->xtvzasu", t, Wr.to_dense())
    H = np.einsum("xtvzasu,bzc->xtvbasuc", t, R.to_dense()).reshape(1600, 1600)
    w0 = np.linalg.eigvalsh(H)[0]
    # the same initial state as the C++ program: srand(12345); Rand() - 0.5
    refkat.srand(12345)
    v = np.array([refkat.Rand() - 0.5 for _ in range(1600)]).reshape(20, 2, 2, 20)
    psi = BSTensor([Din, dout, dout, Dout])
    psi.blocks[(0, 0, 0, 0)] = v
    for (e, iters, nrm), prm in zip(runs, [(1e-9, 200), (1e-16, 20)]):
        assert abs(e - w0) < 1e-8
        o = OLanczos([L, Wl, Wr, R], psi, eff_ham_mul_two_site_state, OParams(*prm))
        assert iters == o.iters
        assert abs(e - o.gs_eng) <= 1e-12 * abs(o.gs_eng)
        assert abs(nrm - 1.0) < 1e-12
    # DMRG-form overload on the same operator: same energy and iteration count as the first case
    assert abs(sb[0][0] - w0) < 1e-8
    assert abs(sb[0][0] - runs[0][0]) <= 1e-12 * abs(w0) and sb[0][1] == runs[0][1]
    # a matvec that is not the layer's own, or one that does not fit the operator count, is refused
    foreign = [l for l in lines if l[0] == "FOREIGN"][0]
    assert foreign[1:] == ["1", "1"]
    tri = [l for l in lines if l[0] == "TRIDIAG"][0]
    e3, v3 = TridiagGsSolver([2.0, 1.0, 3.0], [0.5, 0.25], 3, "V")
    assert abs(float(tri[1]) - e3) < 1e-14 and abs(abs(float(tri[2])) - abs(v3[0])) < 1e-14
    # the three plans drew their workspaces from the context's block cache, and
    # releasing the cache empties it
    cache = [l for l in lines if l[0] == "CACHE"][0]
    mallocs, reuses, cached_after = int(cache[1]), int(cache[2]), int(cache[3])
    assert mallocs > 0 and reuses > 0 and cached_after == 0
