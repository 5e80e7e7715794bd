"""Mesh sizes that are not powers of two (FFTW 2 takes any Nmesh, main.c:1566-1570).

CPU side: the oracle's mixed-radix transform against numpy for lengths with factors 3, 5
and 7, its modal NG sum against the independent regrouped numpy restatement at such a
size, and the C-ABI's size rule (multiples of 16 up to 1024 whose prime factors are 2, 3, 5
and 7; anything else is refused before the device is touched).
"""
import ctypes as C

import numpy as np
import pytest

import aarambam_amd as aar
from aarambam_amd import api
from oracle import pyoracle as orc
from tests import cases, ngref


@pytest.mark.parametrize("N", [12, 14, 20, 24, 30])
def test_oracle_mixed_radix_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    spec = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    ref = np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    np.testing.assert_allclose(orc.c2r(spec), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    real = rng.standard_normal((N, N, N))
    ref2 = np.fft.rfftn(real)
    np.testing.assert_allclose(orc.r2c(real), ref2, rtol=0, atol=1e-12 * np.abs(ref2).max())


@pytest.mark.parametrize("table", ["local", "equilateral"])
def test_oracle_ng_potential_at_24_vs_regrouped_numpy(table):
    N, Nm = 24, 4
    rows = cases.local_rows() if table == "local" else cases.equilateral_rows()
    alpha = orc.alpha_table(Nm, rows)
    x0 = np.zeros((Nm, 2))
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=Nm)
    phig = orc.gaussian(cfg)
    c = orc.cosmology(cfg)
    p = orc.make_params(cfg)
    cc = orc.OrcCosmo(**c)
    buf = np.ascontiguousarray(phig).view(np.float64).reshape(-1).copy()
    rc = orc.lib().orc_ng_potential(C.byref(p), C.byref(cc), np.ascontiguousarray(alpha.reshape(-1)),
                                    np.ascontiguousarray(x0.reshape(-1)), buf, None, None)
    assert rc == 0
    ours = buf.view(np.complex128).reshape(phig.shape)
    ref = ngref.potential(cfg, phig, alpha, x0)
    assert np.linalg.norm(ours - ref) / np.linalg.norm(ref) < 1e-12
    assert np.linalg.norm(ours - phig) / np.linalg.norm(phig) > 1e-6


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_mesh_size_rule_of_the_c_abi():
    ok = [n for n in range(16, 1025, 16) if _smooth7(n)]
    assert 96 in ok and 768 in ok and 1008 in ok and 176 not in ok
    for N in [16, 48, 80, 96, 112, 176, 200, 100, 1040, 8]:
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=4, GlassTileFac=1)
        alpha, x0, _, gbox = cases.job_inputs(cfg, glass_n=4)
        glass = cases.lattice_glass(4, 100.0)
        n = aar.lib().aar_job_bytes(C.byref(api.params_from_dict(cfg)), C.byref(api.tables(alpha, x0, glass, gbox).c),
                                    aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        assert (n > 0) == (N in ok), (N, n)
