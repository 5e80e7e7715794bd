"""CPU checks of the decomposition contractions' oracle (oracle/oracle.c):
against the reference's own tetrapyd.c built from its sources (oracle/_ref,
when present) and against closed forms."""
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
