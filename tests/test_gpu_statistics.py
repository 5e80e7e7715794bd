"""Template check of Phi_NG: binned P(k) and bispectrum on the device.

* The device statistics (aar_job_ng_statistics) against an independent numpy
  evaluation at 16^3: the band-filtered fields built from numpy FFTs, the
  mode functions and regrouped terms of tests/ngref.py, and for the measured
  bispectrum and the triangle counts an explicit enumeration of all
  (k1, k2, k3 = -k1 - k2) on the periodic grid.
* The statistical check at 128^3 (four seeds): the measured power spectrum of
  phi_G against the input P(k), and the measured cross bispectrum
  <phi_G phi_G psi> of the non-Gaussian part against the prediction of the
  input template (local, equilateral and the shipped ScalarI table), bin by
  bin for the bins with many triangles.  The expectation is exact at tree
  level; the tolerances are a few times the scatter over seeds.
"""
import numpy as np
import pytest

from tests import cases, ngref

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def make_job(ctx, cfg, rows, x0=None, prec=aar.AAR_FP64):
    alpha, x0v, glass, gbox = cases.job_inputs(cfg, rows, x0)
    t = aar.tables(alpha, x0v, glass, gbox)
    p = aar.params_from_dict(cfg)
    job = aar.Job(ctx, p, t, prec, aar.AAR_SCHED_BATCHED)
    job.run()
    job.sync()
    return job, alpha, x0v, aar.api.cosmology(p, t)


def numpy_stats(cfg, phig, phi, alpha, x0, cosmo, edges):
    N = cfg["Nmesh"]
    nb = len(edges) - 1
    N3 = float(N) ** 3
    inside, k, k2, _ = ngref.cut_mask(cfg)
    live = inside.copy()
    live[N // 2, :, :] = False
    live[:, N // 2, :] = False
    live[:, :, N // 2] = False
    b = np.digitize(k, edges) - 1
    masks = [live & (b == q) for q in range(nb)]
    real = lambda s: np.fft.irfftn(s, s=(N, N, N), axes=(0, 1, 2)) * N3
    psi = (phi - phig) * N3 / cfg["Fnl"]
    ns = cfg["PrimordialIndex"]
    with np.errstate(divide="ignore", invalid="ignore"):
        P = np.where(live, cosmo["Anorm"] * k ** ns * (cosmo["fac"] * cosmo["Beta"] / k2) ** 2, 0.0)
    Ip = [real(phig * m) for m in masks]
    Is = [real(psi * m) for m in masks]
    J = [real(m.astype(float)) for m in masks]
    n = np.array([(j * j).sum() / N3 for j in J])
    pk = np.array([(i * i).sum() / N3 for i in Ip]) / n
    pth = np.array([(j * real(P * m)).sum() / N3 for j, m in zip(J, masks)]) / n
    terms = ngref.ng_terms(alpha, cfg["N_modes"])
    modes = sorted({t[0] for v in terms.values() for t in v} | {t[1] for v in terms.values() for t in v})
    with np.errstate(divide="ignore", invalid="ignore"):
        q = {m: np.where(live, ngref.weights(cfg, x0, k, k2, m) * k ** (4 - ns), 0.0) for m in modes}
        gK = {K: np.where(live, ngref.weights(cfg, x0, k, k2, K), 0.0) for K in terms}
    Q = {m: [real(P * q[m] * mk) for mk in masks] for m in modes}
    G = {K: [real(gK[K] * mk) for mk in masks] for K in terms}
    bm = np.zeros((nb, nb, nb))
    bt = np.zeros((nb, nb, nb))
    nt = np.zeros((nb, nb, nb))
    for b1 in range(nb):
        for b2 in range(nb):
            for b3 in range(nb):
                bm[b1, b2, b3] = (Ip[b1] * Ip[b2] * Is[b3]).sum()
                nt[b1, b2, b3] = (J[b1] * J[b2] * J[b3]).sum() / N3
                s = 0.0
                for K, lst in terms.items():
                    for i, j, c in lst:
                        s += c * ((Q[i][b1] * Q[j][b2] + Q[j][b1] * Q[i][b2]) * G[K][b3]).sum()
                bt[b1, b2, b3] = N3 * s
    return dict(pk_meas=pk, pk_theory=pth, nmodes=n, b_meas=bm, b_theory=bt, ntri=nt), (masks, phig, psi)


def enumerate_triangles(cfg, edges, phig, psi):
    """N^3 sum over k1, k2 of phi(k1) phi(k2) psi(k3), k3 = -k1 - k2 (mod N), per bin triple,
    on the full periodic grid (cube cut at Nsample = Nmesh: only the Nyquist planes and the
    origin are excluded)."""
    N = cfg["Nmesh"]
    assert cfg["Nsample"] == N and cfg.get("SphereMode", 0) == 0
    full = lambda s: np.fft.fftn(np.fft.irfftn(s, s=(N, N, N), axes=(0, 1, 2)))
    ph, ps = full(phig).reshape(-1), full(psi).reshape(-1)
    idx = np.arange(N)
    kv = np.where(idx < N // 2, idx, -(N - idx)) * 2 * np.pi / cfg["Box"]
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    KX, KY, KZ = np.meshgrid(kv, kv, kv, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    live = ~((X == 0) & (Y == 0) & (Z == 0)) & (X != N // 2) & (Y != N // 2) & (Z != N // 2)
    bins = np.where(live, np.digitize(kk, edges) - 1, -1)
    bins[bins >= len(edges) - 1] = -1
    bins = bins.reshape(-1)
    nb = len(edges) - 1
    B = np.zeros((nb, nb, nb), complex)
    T = np.zeros((nb, nb, nb))
    ks = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)
    for a in np.nonzero(bins >= 0)[0]:
        k3 = (-(ks[a][None, :] + ks)) % N
        i3 = (k3[:, 0] * N + k3[:, 1]) * N + k3[:, 2]
        b2, b3 = bins, bins[i3]
        ok = (b2 >= 0) & (b3 >= 0)
        np.add.at(B, (bins[a], b2[ok], b3[ok]), ph[a] * ph[ok] * ps[i3[ok]])
        np.add.at(T, (bins[a], b2[ok], b3[ok]), 1.0)
    return float(N) ** 3 * B.real, T


def test_statistics_match_numpy_and_enumeration(ctx):
    cfg = cases.base_config(Nmesh=16, Nsample=16, Nmax=16, N_modes=4, Box=200.0, Seed=9)
    job, alpha, x0, cosmo = make_job(ctx, cfg, cases.equilateral_rows())
    kf = 2 * np.pi / cfg["Box"]
    edges = kf * np.array([1.0, 2.5, 4.0, 5.5, 7.5])
    dev = job.ng_statistics(edges)
    ref, (_, phig, psi) = numpy_stats(cfg, job.field("phig_spec"), job.field("phi_spec"), alpha, x0, cosmo,
                                      edges)
    job.close()
    for key in ["pk_meas", "pk_theory", "nmodes", "ntri"]:
        np.testing.assert_allclose(dev[key], ref[key], rtol=1e-9, atol=1e-9 * np.abs(ref[key]).max())
    for key in ["b_meas", "b_theory"]:
        np.testing.assert_allclose(dev[key], ref[key], rtol=0, atol=1e-9 * np.abs(ref[key]).max())
    B, T = enumerate_triangles(cfg, edges, phig, psi)
    np.testing.assert_allclose(dev["ntri"], T, rtol=0, atol=1e-6)
    np.testing.assert_allclose(dev["b_meas"], B, rtol=0, atol=1e-9 * np.abs(B).max())


@pytest.mark.parametrize("tpl", ["local", "equilateral", "scalari"])
def test_phi_ng_matches_its_template(ctx, tpl):
    N, L = 128, 1000.0
    kf = 2 * np.pi / L
    edges = kf * np.array([1, 4, 8, 14, 22, 32, 44, 60.5])
    acc = []
    for seed in [1, 2, 3, 4]:
        if tpl == "scalari":
            cfg, rows, x0 = cases.shipped_table_config(N, Seed=seed, Box=L)
        else:
            cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Box=L, GlassTileFac=2, Seed=seed)
            rows, x0 = (cases.local_rows() if tpl == "local" else cases.equilateral_rows()), None
        job, *_ = make_job(ctx, cfg, rows, x0, aar.AAR_FP32)
        r = job.ng_statistics(edges)
        job.close()
        big = r["nmodes"] > 3e4  # P(k): 1 / sqrt(n) scatter below 0.6 %
        assert np.all(np.abs(r["pk_meas"][big] / r["pk_theory"][big] - 1) < 0.03), r["pk_meas"] / r["pk_theory"]
        acc.append(r)
    bm = np.mean([r["b_meas"] for r in acc], axis=0)
    bt = acc[0]["b_theory"]  # the expectation does not depend on the seed
    nt = acc[0]["ntri"]
    sel = nt > 2e9
    ratio = bm[sel] / bt[sel]
    print(tpl, "meas/theory over", sel.sum(), "bin triples:", np.round(np.sort(ratio), 3))
    assert np.all(np.abs(ratio - 1) < 0.25), ratio
    assert abs(np.average(ratio, weights=nt[sel]) - 1) < 0.03, np.average(ratio, weights=nt[sel])


def test_templates_are_distinguished(ctx):
    """The local and equilateral predictions differ in shape by far more than the tolerance above."""
    N, L = 64, 1000.0
    kf = 2 * np.pi / L
    edges = kf * np.array([1, 4, 8, 14, 22, 30.5])
    shapes = []
    for rows in [cases.local_rows(), cases.equilateral_rows()]:
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Box=L, Seed=1)
        job, *_ = make_job(ctx, cfg, rows, None, aar.AAR_FP32)
        r = job.ng_statistics(edges)
        job.close()
        sel = r["ntri"] > 1e7
        v = r["b_theory"][sel] / r["ntri"][sel]
        shapes.append(v / np.linalg.norm(v))
    # bin-averaged shapes as unit vectors: far from parallel (the equilateral one changes sign)
    assert abs(float(np.dot(shapes[0], shapes[1]))) < 0.9
