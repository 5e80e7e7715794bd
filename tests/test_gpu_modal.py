"""Decomposition contractions on the device (aar_mode_bispectra_covariance,
aar_mode_QS) against the CPU oracle, which equals the reference's own
tetrapyd.c bit for bit (tests/test_modal.py).  The device sums over the
tetrapyd points in a different order (MFMA tiles, split over blocks), so the
bar is 1e-12 relative to the largest entry."""
import os

import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases, modal

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("p_max,nk", [(4, 16), (8, 40), (12, 64)])
def test_covariance_matches_oracle(ctx, p_max, nk):
    g = modal.basis_grid(p_max=p_max, k_grid_size=nk)
    dev = ctx.mode_covariance(g)
    ref = orc.mode_covariance(g)
    scale = np.abs(ref).max()
    assert np.abs(dev - ref).max() <= 1e-12 * scale, np.abs(dev - ref).max() / scale
    if orc.ref_lib() is not None and p_max <= 8:
        ref2 = orc.mode_covariance(g, use_ref=True)
        assert np.abs(dev - ref2).max() <= 1e-12 * scale


def test_qs_matches_oracle(ctx):
    g = modal.basis_grid(p_max=10, k_grid_size=48)
    S = np.random.default_rng(11).standard_normal((3, len(g["w"])))
    dev = ctx.mode_qs(g, S)
    ref = orc.mode_qs(g, S)
    assert np.abs(dev - ref).max() <= 1e-12 * np.abs(ref).max()


def test_out_of_range_indices_are_rejected(ctx):
    g = modal.basis_grid(p_max=4, k_grid_size=16)
    bad = dict(g)
    bad["i2"] = g["i2"].copy()
    bad["i2"][5] = 16  # == k_npts
    with pytest.raises(aar.AarError):
        ctx.mode_covariance(bad)
    bad = dict(g)
    bad["p3"] = g["p3"].copy()
    bad["p3"][0] = -1
    with pytest.raises(aar.AarError):
        ctx.mode_qs(bad, np.ones((1, len(g["w"]))))


# ---------------------------------------------------------------- basis expansion
def _shape(grid, kind):
    from oracle import pymodal
    k = grid["k"]
    k1, k2, k3 = k[grid["i1"]], k[grid["i2"]], k[grid["i3"]]
    if kind == "local":
        return ((k1 * k1 / (k2 * k3) + k2 * k2 / (k1 * k3) + k3 * k3 / (k1 * k2)) / 3)[None, :]
    if kind == "equilateral":
        return ((k1 + k2 - k3) * (k2 + k3 - k1) * (k3 + k1 - k2) / (k1 * k2 * k3))[None, :]
    if kind == "scalari":  # the collider template of the benchmark's table (mass 1)
        from tests import templates
        return templates.shape(grid, lambda a, b, c: templates.scalar_i(a, b, c, 0.9660499, 1.0))
    if kind == "two":  # the second shape's selection is driven by the first one's residual, as in the reference
        return np.concatenate([_shape(grid, "local"), _shape(grid, "equilateral")])
    Q = pymodal.mode_bispectra(grid)
    n = int(np.nonzero((grid["p1"] == 5) & (grid["p2"] == 2) & (grid["p3"] == 1))[0][0])
    return Q[n][None, :]


@pytest.mark.parametrize("kind,kw", [
    ("local", dict(RR_tol=1e-8, max_modes=40)),
    ("equilateral", dict(RR_tol=1e-8, max_modes=40, rtol=1e-10, Niter=1e3)),
    ("mode", dict(RR_tol=1e-12, cond_tol=1e14, max_modes=60)),
    ("local", dict(max_modes=-1)),
    ("two", dict(RR_tol=1e-8, max_modes=30, rtol=1e-10, Niter=1e3, bad_modes="decompose")),
    ("scalari", dict(RR_tol=1e-4, max_modes=60, rtol=1e-10, Niter=1e3, bad_modes="decompose")),
])
def test_basis_expansion_matches_oracle(ctx, kind, kw):
    """Same modes selected as by the oracle's restatement of Basis.basis_expansion, coefficients
    to 1e-6 of the largest, convergence figures to 1e-9.  With every mode selected
    (max_modes -1) the Gram matrix's condition number is ~1e15 and the coefficients of two
    conjugate-gradient runs differ in rounding; there the fitted shapes are compared."""
    from oracle import pymodal
    every = kw.get("max_modes") == -1
    g = modal.basis_grid(p_max=6 if every else 8, k_grid_size=30)
    if kw.get("bad_modes") == "decompose":  # the mask of Aarambam/basis/decompose.py:125-129
        p1, p2, p3 = g["p1"], g["p2"], g["p3"]
        kw = dict(kw, bad_modes=(p1 < 4) & ~(((p1 == 1) & (p2 == 1) & (p3 == 1)) | ((p1 == 2) & (p2 == 1) & (p3 == 0))
                                             | ((p1 == 3) & (p2 == 0) & (p3 == 0))))
    S = _shape(g, kind)
    print("expansion", kind, kw, len(g["p1"]), "modes", len(g["w"]), "points", flush=True)
    ra, rcov, rcorr, rmse, rmask = pymodal.basis_expansion(g, S, **kw)
    da, dcov, dcorr, dmse, dmask = ctx.mode_basis_expansion(g, S, **kw)
    np.testing.assert_array_equal(dmask, rmask)
    np.testing.assert_allclose(dcov, rcov, rtol=1e-12)
    if every:
        Q = pymodal.mode_bispectra(g)
        np.testing.assert_allclose(da @ Q, ra @ Q, rtol=0, atol=1e-6 * np.abs(S).max())
        np.testing.assert_allclose(dmse, rmse, rtol=1e-3)
        return
    np.testing.assert_allclose(da, ra, rtol=0, atol=1e-6 * np.abs(ra).max())
    np.testing.assert_allclose(dcorr, rcorr, rtol=0, atol=1e-9)
    np.testing.assert_allclose(dmse, rmse, rtol=1e-4, atol=1e-12)


def test_basis_expansion_rejects_bad_input(ctx):
    g = modal.basis_grid(p_max=4, k_grid_size=12)
    S = np.ones((1, len(g["w"])))
    bad = dict(g, i1=np.where(np.arange(len(g["i1"])) == 3, 12, g["i1"]))
    with pytest.raises(aar.AarError):
        ctx.mode_basis_expansion(bad, S)
    with pytest.raises(aar.AarError):
        ctx.mode_basis_expansion(g, S, max_modes=-2)


def test_decomposition_tables_drive_an_ic_job(ctx, tmp_path):
    """Template -> aarambam_amd.decompose (expansion on the device) -> AlphaTable.dat / AiTable.dat in
    the reference's format -> aar_load_alpha / aar_load_x0 -> an FP64 IC job that matches the oracle
    run on the same tables to 1e-10."""
    import ctypes as C
    from aarambam_amd import decompose
    from tests import templates
    from tests.test_gpu_parity import _compare_job
    ns, L, N = 0.9649, 1000.0, 32
    r = decompose.decompose(ctx, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0), 8, ns, L, N, 1e-4, 30,
                            k_grid_size=24)
    ref = pymodal_expansion_check(r, ns)
    assert ref
    decompose.write_tables(r, str(tmp_path))
    alpha = np.empty(8 ** 3)
    x0 = np.empty(16)
    dp = C.POINTER(C.c_double)
    assert aar.lib().aar_load_alpha(str(tmp_path / "AlphaTable.dat").encode(), 8, alpha.ctypes.data_as(dp)) == 0
    assert aar.lib().aar_load_x0(str(tmp_path / "AiTable.dat").encode(), 8, x0.ctypes.data_as(dp)) == 0
    rows = r["AlphaTable"][r["AlphaTable"][:, 3] != 0]
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=8, Box=L, PrimordialIndex=ns)
    np.testing.assert_array_equal(alpha.reshape(8, 8, 8), orc.alpha_table(8, rows))
    _compare_job(ctx, cfg, rows, x0.reshape(8, 2), aar.AAR_FP64, aar.AAR_SCHED_BATCHED, 1e-10)


def pymodal_expansion_check(r, ns):
    """The device expansion inside decompose() selects what the oracle's restatement selects."""
    from aarambam_amd import decompose
    from oracle import pymodal
    from tests import templates
    g = r["grid"]
    S = decompose.shape_values(g, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0))[None, :]
    ra = pymodal.basis_expansion(g, S, rtol=1e-10, Niter=1e3, bad_modes=decompose.dropped_modes(g), RR_tol=1e-4,
                                 max_modes=30)
    np.testing.assert_array_equal(ra[4][0], r["selected"])
    np.testing.assert_allclose(r["AlphaTable"][:, 3], ra[0][0], rtol=0, atol=1e-6 * np.abs(ra[0][0]).max())
    return True


GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _sym_cases(rng):
    """Symmetric test matrices for the condition-number kernel: normalised Gram matrices (the
    expansion's kind), shifted to be indefinite, block-diagonal (columns already reduced),
    diagonal, and sizes either side of one block's 256 threads."""
    out = []
    for m in (1, 2, 3, 17, 64, 130, 257, 300):
        X = rng.standard_normal((m + 7, m)) * np.exp(rng.uniform(-4, 0, m))
        G = X.T @ X
        dn = np.sqrt(np.diag(G))
        G = G / dn[:, None] / dn[None, :]
        out.append(("gram", G))
        out.append(("indefinite", G - 0.5 * np.eye(m)))
        if m >= 3:
            B = G.copy()
            B[: m // 2, m // 2:] = 0
            B[m // 2:, : m // 2] = 0
            out.append(("block", B))
        out.append(("diagonal", np.diag(rng.uniform(-2, 3, m))))
    return out


def test_condition_kernel_matches_numpy(ctx):
    """The extreme eigenvalue magnitudes behind the expansion's condition numbers against
    numpy's eigvalsh: within 1e-13 m of the largest magnitude (the Householder reduction's
    backward error), and np.linalg.cond to 1e-9 where it is below 1e4."""
    rng = np.random.default_rng(77)
    for kind, A in _sym_cases(rng):
        m = A.shape[0]
        ext = ctx.sym_eig_extremes(np.stack([A, A.T.copy()]))
        ev = np.abs(np.linalg.eigvalsh(A))
        tol = 1e-13 * max(m, 4) * ev.max()
        for e in ext:
            assert abs(e[0] - ev.min()) <= tol and abs(e[1] - ev.max()) <= tol, (kind, m, e, ev.min(), ev.max())
            c = np.linalg.cond(A)
            if c < 1e4:
                assert abs(e[1] / e[0] / c - 1) <= 1e-9, (kind, m, e[1] / e[0], c)



@pytest.mark.parametrize("name", ["equil", "scalari"])
def test_device_expansion_matches_reference_vectors(ctx, name):
    """The device expansion against the reference's basis_expansion (vectors written by
    tests/golden/make_expansion_vectors.py from the reference's cmbbest, 40-point k grid,
    16 modes, decompose.py's settings): the same modes picked, coefficients to 1e-6 of the
    largest, correlation to 1e-9."""
    f = np.load(os.path.join(GOLD, "expansion_grid40.npz"))
    g = {k: f[k] for k in ["k", "w", "i1", "i2", "i3", "evals", "p1", "p2", "p3"]}
    corr, mse, tol, mm = f[f"corr_{name}"]
    alpha, _, c, m, sel = ctx.mode_basis_expansion(g, f[f"S_{name}"][None, :], rtol=1e-10, Niter=1e3,
                                                   bad_modes=f["bad_modes"], RR_tol=tol, max_modes=int(mm))
    ref = f[f"alpha_{name}"]
    np.testing.assert_array_equal(alpha[0] != 0, ref != 0)
    np.testing.assert_allclose(alpha[0], ref, rtol=0, atol=1e-6 * np.abs(ref).max())
    np.testing.assert_allclose(c[0], corr, rtol=0, atol=1e-9)


def test_scalari16_decomposition_matches_reference_table(ctx):
    """aarambam_amd.decompose (150-point k grid, expansion on the device) for the benchmark's
    template -- ScalarI mass 1, 16 modes, n_s 0.9649, L 2000, Nmax 512, ModeTol 1e-4,
    MaxModeCount 150 -- against the AlphaTable the reference's
    BasicBasisDecompose.generate_alpha writes for it: the same 132 nonzero rows, coefficients to
    1e-6 of the largest, the same correlation to 1e-9."""
    from aarambam_amd import decompose
    from tests import templates
    path = os.path.join(GOLD, "scalari16_alphatable.txt")
    ref = np.loadtxt(path)
    with open(path) as fh:
        corr = float(fh.readline().split("correlation ")[1].split(";")[0])
    r = decompose.decompose(ctx, lambda a, b, c: templates.scalar_i(a, b, c, 0.9649, 1.0), 16, 0.9649, 2000.0, 512,
                            1e-4, 150)
    t = r["AlphaTable"][r["AlphaTable"][:, 3] != 0]
    np.testing.assert_array_equal(t[:, :3], ref[:, :3])
    np.testing.assert_allclose(t[:, 3], ref[:, 3], rtol=0, atol=1e-6 * np.abs(ref[:, 3]).max())
    assert abs(r["correlation"] - corr) < 1e-9
