"""CPU checks of the decomposition contractions' oracle (oracle/oracle.c):
against the reference's own tetrapyd.c built from its sources (oracle/_ref,
when present) and against closed forms."""
import os

import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import modal


@pytest.fixture(scope="module")
def grid():
    return modal.basis_grid(p_max=6, k_grid_size=18)


def test_oracle_matches_reference_build(grid):
    if orc.ref_lib() is None:
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    cov = orc.mode_covariance(grid)
    ref = orc.mode_covariance(grid, use_ref=True)
    np.testing.assert_array_equal(cov, ref)  # same loops, same operation order
    S = np.random.default_rng(3).standard_normal((3, len(grid["w"])))
    np.testing.assert_array_equal(orc.mode_qs(grid, S), orc.mode_qs(grid, S, use_ref=True))


def test_constant_mode_functions_closed_form(grid):
    g = dict(grid)
    a = np.array([1.5, -0.5, 2.0, 0.25, 3.0, -1.25])
    g["evals"] = np.repeat(a[:, None], grid["evals"].shape[1], axis=1)
    cov = orc.mode_covariance(g)
    q = a[g["p1"]] * a[g["p2"]] * a[g["p3"]]  # Q_n = a_p1 a_p2 a_p3 at every point
    np.testing.assert_allclose(cov, np.outer(q, q) * g["w"].sum(), rtol=1e-13)
    qs = orc.mode_qs(g, np.ones((1, len(g["w"]))))
    np.testing.assert_allclose(qs[0], q * g["w"].sum(), rtol=1e-13)


def test_covariance_is_a_gram_matrix(grid):
    """cov = (w Q) Q^T: symmetric, positive semi-definite, diagonal = sum w Q_n^2 (numpy, independent)."""
    cov = orc.mode_covariance(grid)
    ev = grid["evals"]
    Q = np.array([(ev[a][grid["i1"]] * ev[b][grid["i2"]] * ev[c][grid["i3"]]
                   + ev[a][grid["i1"]] * ev[b][grid["i3"]] * ev[c][grid["i2"]]
                   + ev[a][grid["i2"]] * ev[b][grid["i1"]] * ev[c][grid["i3"]]
                   + ev[a][grid["i2"]] * ev[b][grid["i3"]] * ev[c][grid["i1"]]
                   + ev[a][grid["i3"]] * ev[b][grid["i1"]] * ev[c][grid["i2"]]
                   + ev[a][grid["i3"]] * ev[b][grid["i2"]] * ev[c][grid["i1"]]) / 6
                  for a, b, c in zip(grid["p1"], grid["p2"], grid["p3"])])
    np.testing.assert_allclose(cov, (Q * grid["w"]) @ Q.T, rtol=1e-10, atol=1e-14 * np.abs(cov).max())
    assert np.linalg.eigvalsh((cov + cov.T) / 2).min() > -1e-10 * np.abs(cov).max()


# ---------------------------------------------------------------- basis expansion (oracle/pymodal.py)
def _q_matrix(grid):
    from oracle import pymodal
    return pymodal.mode_bispectra(grid)


def _good(grid):
    p1, p2, p3 = grid["p1"], grid["p2"], grid["p3"]
    return ~(((p1 <= 2) & (p2 == 0) & (p3 == 0)) | ((p1 == 1) & (p2 == 1) & (p3 == 0)))


def _index(grid, t):
    return int(np.nonzero((grid["p1"] == t[0]) & (grid["p2"] == t[1]) & (grid["p3"] == t[2]))[0][0])


def test_expansion_of_a_shape_in_the_starting_span(grid):
    """S = sum of four starting modes: no selection step, coefficients recovered, correlation 1."""
    from oracle import pymodal
    Q = _q_matrix(grid)
    c = {(3, 0, 0): 1.0, (2, 1, 1): -0.5, (3, 2, 1): 0.25, (1, 1, 1): 2.0}
    S = sum(v * Q[_index(grid, t)] for t, v in c.items())[None, :]
    alpha, cov, corr, mse, masks = pymodal.basis_expansion(grid, S, RR_tol=1e-14)
    start = (grid["p1"] <= 3) & (grid["p2"] <= 3) & (grid["p3"] <= 3)
    np.testing.assert_array_equal(masks[0], start)
    expect = np.zeros(len(grid["p1"]))
    for t, v in c.items():
        expect[_index(grid, t)] = v
    np.testing.assert_allclose(alpha[0], expect, atol=1e-7 * np.abs(expect).max())
    assert abs(corr[0] - 1) < 1e-12 and mse[0] < 1e-12
    np.testing.assert_allclose(cov[0, 0], (S[0] * grid["w"]) @ S[0], rtol=1e-14)


def test_expansion_over_all_modes_is_the_weighted_least_squares_fit(grid):
    """max_modes -1 selects every mode: alpha is the w-weighted least-squares fit of S by the
    modes outside the dropped set (numpy lstsq, independent of the CG path)."""
    from oracle import pymodal
    Q = _q_matrix(grid)
    w = grid["w"]
    k = grid["k"]
    k1, k2, k3 = k[grid["i1"]], k[grid["i2"]], k[grid["i3"]]
    S = ((k1 * k1 / (k2 * k3) + k2 * k2 / (k1 * k3) + k3 * k3 / (k1 * k2)) / 3)[None, :]
    alpha, _, corr, mse, masks = pymodal.basis_expansion(grid, S, max_modes=-1)
    assert masks[0].all()
    good = _good(grid)
    A = (Q[good] * np.sqrt(w)).T
    sol, *_ = np.linalg.lstsq(A, S[0] * np.sqrt(w), rcond=None)
    rec, ref = alpha[0] @ Q, sol @ Q[good]
    np.testing.assert_allclose(rec, ref, atol=1e-7 * np.abs(S).max())
    assert np.all(alpha[0][~good] == 0)
    resid = ((S[0] - ref) ** 2 * w).sum() / ((S[0] ** 2) * w).sum()
    assert mse[0] == pytest.approx(resid, rel=1e-4)  # MSE from the cancelling form SS + RR - 2 SR


def test_selection_loop_recovers_a_mode_outside_the_start(grid):
    """S = one mode with p1 = 4: the loop keeps adding modes while the residual is above the
    tolerance, so it ends with that mode selected and the shape reproduced (the coefficients of the
    other, nearly collinear modes are zero to the conditioning of their Gram matrix)."""
    from oracle import pymodal
    Q = _q_matrix(grid)
    b = _index(grid, (4, 2, 1))
    alpha, _, corr, mse, masks = pymodal.basis_expansion(grid, Q[b][None, :], RR_tol=1e-12, cond_tol=1e14)
    assert masks[0][b] and masks[0].sum() > 20
    assert alpha[0][b] == pytest.approx(1.0, abs=1e-4)
    assert np.abs(np.delete(alpha[0], b)).max() < 1e-4
    assert mse[0] < 1e-12 and abs(corr[0] - 1) < 1e-12


def test_scalar_i_template_is_normalised_at_the_equilateral_point():
    """The Scalar-I shape is normalised so that S(k, k, k) = k^(2(n_s - 1)) (each k enters as
    k^((4 - n_s)/3) and every term is homogeneous of degree 0), and it is symmetric."""
    from tests import templates
    ns = 0.9649
    k = np.geomspace(1e-3, 1.0, 7)
    g = {"k": k, "i1": np.arange(7), "i2": np.arange(7), "i3": np.arange(7)}
    S = templates.shape(g, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0))[0]
    np.testing.assert_allclose(S, k ** (2 * (ns - 1)), rtol=1e-12)
    g2 = {"k": k, "i1": np.array([6, 5, 4]), "i2": np.array([4, 6, 5]), "i3": np.array([5, 4, 6])}
    S2 = templates.shape(g2, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0))[0]
    np.testing.assert_allclose(S2, S2[0], rtol=1e-13)


def test_product_basis_grid_matches_the_test_restatement():
    """aarambam_amd.decompose.basis_grid (product) and tests/modal.basis_grid (test input) restate
    cmbbest.Basis independently; the Legendre modes differ only in evaluation (power series of
    scipy.special.legendre in the product, as the reference, recurrence in the tests)."""
    from aarambam_amd import decompose
    kmin, kmax, ns = 2 * np.pi / 1000, 2 * np.pi / 1000 * 256, 0.9660499
    a = decompose.basis_grid(12, ns, kmin, kmax, 40)
    b = modal.basis_grid(p_max=12, k_grid_size=40, k_min=kmin, k_max=kmax, n_s=ns)
    for key in ["i1", "i2", "i3", "p1", "p2", "p3"]:
        np.testing.assert_array_equal(a[key], b[key])
    np.testing.assert_allclose(a["w"], b["w"], rtol=1e-13)
    np.testing.assert_allclose(a["evals"], b["evals"], rtol=0, atol=1e-11 * np.abs(b["evals"]).max())
    # AiTable: x0 of the reference's shipped table (defaults/AiTable.dat: N_modes 15, Nmax 256, L 512)
    from tests import cases
    ai = decompose.ai_table(15, ns, 2 * np.pi / 512, 2 * np.pi / 512 * 256)
    np.testing.assert_allclose(ai[:, 1], cases.x0_file(15)[:, 0], rtol=1e-14)


GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold40():
    """Vectors written by tests/golden/make_expansion_vectors.py from the reference's own cmbbest
    (Basis grid, shapes and basis_expansion on a 40-point k grid, 16 modes)."""
    return dict(np.load(os.path.join(GOLD, "expansion_grid40.npz")))


def test_basis_grid_matches_reference_vectors(gold40):
    """decompose.basis_grid against cmbbest.Basis itself: the k grid, tetrapyd points and weights,
    mode functions and 3-D mode indices, bit for bit."""
    from aarambam_amd import decompose
    ns, kmin, kmax = gold40["params"]
    g = decompose.basis_grid(16, ns, kmin, kmax, 40)
    for key in ["k", "w", "i1", "i2", "i3", "evals", "p1", "p2", "p3"]:
        np.testing.assert_array_equal(g[key], gold40[key], err_msg=key)


def test_scalar_i_shape_matches_reference_vectors(gold40):
    """The ScalarI (mass 1) shape on the tetrapyd points against the reference's model
    (Aarambam/models/Colliders.py with decompose.py's symmetrisation)."""
    from aarambam_amd import decompose
    from tests import templates
    ns, kmin, kmax = gold40["params"]
    g = decompose.basis_grid(16, ns, kmin, kmax, 40)
    S = templates.shape(g, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0))[0]
    np.testing.assert_allclose(S, gold40["S_scalari"], rtol=0, atol=1e-13 * np.abs(gold40["S_scalari"]).max())


@pytest.mark.parametrize("name", ["equil", "scalari"])
def test_oracle_expansion_matches_reference_vectors(gold40, name):
    """oracle/pymodal.basis_expansion against the reference's basis_expansion with decompose.py's
    settings: the same modes (the OMP loop's picks, 132 for ScalarI), the same coefficients to
    1e-12 of the largest, the same correlation and MSE.  The ScalarI case walks ~115 selection
    steps (about a minute)."""
    from oracle import pymodal
    g = {k: gold40[k] for k in ["k", "w", "i1", "i2", "i3", "evals", "p1", "p2", "p3"]}
    corr, mse, tol, mm = gold40[f"corr_{name}"]
    alpha, _, c, m, _ = pymodal.basis_expansion(g, gold40[f"S_{name}"][None, :], rtol=1e-10, Niter=1e3,
                                                bad_modes=gold40["bad_modes"], RR_tol=tol, max_modes=int(mm))
    ref = gold40[f"alpha_{name}"]
    np.testing.assert_array_equal(alpha[0] != 0, ref != 0)
    np.testing.assert_allclose(alpha[0], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(c[0], corr, rtol=0, atol=1e-12)
    np.testing.assert_allclose(m[0], mse, rtol=1e-9)
