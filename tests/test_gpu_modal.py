"""Decomposition contractions on the device (aar_mode_bispectra_covariance,
aar_mode_QS) against the CPU oracle, which equals the reference's own
tetrapyd.c bit for bit (tests/test_modal.py).  The device sums over the
tetrapyd points in a different order (MFMA tiles, split over blocks), so the
bar is 1e-12 relative to the largest entry."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import modal

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
