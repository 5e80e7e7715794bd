"""Pins the CPU oracle against known answers, closed forms and independent restatements."""
import numpy as np
import pytest
from scipy import integrate, special

from oracle import pyoracle as orc
from tests import cases, ngref


def test_ranlxd_gsl_known_answers():
    # GSL rng/test.c: rng_test(gsl_rng_ranlxd2, 1, 10000, 3949287736).  ranlxd1 runs
    # the same code with luxury p=202 instead of 397 (Luescher's level 1).
    assert orc.ranlxd_nth_get(1, 397, 10000) == 3949287736
    assert orc.ranlxd_nth_get(1, 202, 10000) == 1998227290
    # seed 0 is mapped to 1 by GSL's ranlxd_set
    assert orc.ranlxd_nth_get(0, 202, 100) == orc.ranlxd_nth_get(1, 202, 100)


def test_ranlxd_draws_are_48bit_fractions():
    u = orc.ranlxd_draws(12345, 5000)
    assert np.all((u >= 0) & (u < 1))
    m = u * 2.0 ** 48
    assert np.all(m == np.floor(m))
    assert abs(u.mean() - 0.5) < 0.02


def test_legendre_matches_scipy():
    xs = np.linspace(-1, 1, 101)
    for l in range(0, 20):
        ours = np.array([orc.legendre(l, x) for x in xs])
        np.testing.assert_allclose(ours, special.eval_legendre(l, xs), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_seedtable_shell_order(N):
    st = orc.seedtable(N, 777)
    big = orc.seedtable(2 * N, 777)
    # main.c:321-349 fills shells i = 0..N/2-1 in a fixed order, so the low-k seeds
    # (indices measured from both ends of each axis) do not depend on N.
    h = N // 2
    lo = np.r_[0:h]
    np.testing.assert_array_equal(st[np.ix_(lo, lo)], big[np.ix_(lo, lo)])
    np.testing.assert_array_equal(st[np.ix_(N - 1 - lo, lo)], big[np.ix_(2 * N - 1 - lo, lo)])
    np.testing.assert_array_equal(st[np.ix_(lo, N - 1 - lo)], big[np.ix_(lo, 2 * N - 1 - lo)])
    # every seed is a 31-bit draw; no entry left unwritten (a zero seed would be 1 in 2^31)
    assert np.all(st > 0) and np.all(st < 2 ** 31)


def test_seedtable_first_draws():
    st = orc.seedtable(4, 42)
    u = orc.ranlxd_draws(42, 16)
    seeds = (0x7FFFFFFF * u).astype(np.uint32)
    # shell 0: (0,0)=first draw [j<i+1 loop], then (3,0), (0,3), (3,3)
    assert st[0, 0] == seeds[0] and st[3, 0] == seeds[1] and st[0, 3] == seeds[2] and st[3, 3] == seeds[3]


@pytest.mark.parametrize("N", [8, 16])
def test_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    spec = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    # arbitrary (non-Hermitian) input: numpy's irfftn ignores the same imaginary parts as FFTW
    ref = np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    np.testing.assert_allclose(orc.c2r(spec), ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    real = rng.standard_normal((N, N, N))
    ref2 = np.fft.rfftn(real)
    np.testing.assert_allclose(orc.r2c(real), ref2, rtol=0, atol=1e-12 * np.abs(ref2).max())


def test_growth_einstein_de_sitter_closed_form():
    cfg = cases.base_config(Omega=1.0, OmegaLambda=0.0)
    # growth(a) = H(a) int_0^a (a'/H(a')... )^1.5 = a/2.5 for Omega=1 (power.c:585-598)
    for a in [1e-3, 0.02, 0.5, 1.0]:
        # qromb stops on its own error estimate (EPS 1e-8, nrsrc/qromb.c:2); the a^1.5
        # integrand converges slowly near 0, so the true error is a few 1e-7.
        assert orc.lib().orc_qromb_growth(orc.make_params(cfg), a) == pytest.approx(a / 2.5, rel=5e-7)
    c = orc.cosmology(cfg)
    assert c["D0"] == pytest.approx(1.0, rel=1e-6)
    assert c["Dplus"] == pytest.approx(1 + cfg["Redshift"], rel=1e-6)


def test_growth_lcdm_vs_quad():
    cfg = cases.base_config()
    Om, OL = cfg["Omega"], cfg["OmegaLambda"]
    f = lambda a: (a / (Om + (1 - Om - OL) * a + OL * a ** 3)) ** 1.5
    for a in [0.02, 0.3, 1.0]:
        H = np.sqrt(Om / a ** 3 + (1 - Om - OL) / a ** 2 + OL)
        ref = H * integrate.quad(f, 0, a, epsabs=0, epsrel=1e-12)[0]
        assert orc.lib().orc_qromb_growth(orc.make_params(cfg), a) == pytest.approx(ref, rel=5e-7)


def test_sigma2_power_law_vs_quad():
    # WhichTransfer 0 gives T=1 (tk_Efstathiou returns 1), so P = Norm k^ns
    cfg = cases.base_config(WhichTransfer=0)
    R = 8.0
    ns = cfg["PrimordialIndex"]

    def integrand(k):
        kr = R * k
        w = 3 * (np.sin(kr) / kr ** 3 - np.cos(kr) / kr ** 2)
        return 4 * np.pi * k * k * w * w * k ** ns

    ref = integrate.quad(integrand, 1e-12, 500.0 / R, limit=2000, epsrel=1e-11)[0]
    ours = orc.lib().orc_sigma2_tophat(orc.make_params(cfg), 1.0, R)
    assert ours == pytest.approx(ref, rel=1e-7)


def test_tabulated_transfer_nodes():
    tk, tt = cases.load_transfer()
    cfg = cases.base_config(WhichTransfer=2, transfer_k=tk, transfer_t=tt)
    for i in [0, 5, 50, 100, len(tk) - 1]:
        assert orc.transfer(cfg, tk[i]) == pytest.approx(tt[i] / tt[0], rel=1e-12)
    assert orc.transfer(cfg, tk[-1] * 1.01) == 0.0
    # log-log interpolation midway between two nodes
    km = np.sqrt(tk[40] * tk[41])
    assert orc.transfer(cfg, km) == pytest.approx(np.sqrt(tt[40] * tt[41]) / tt[0], rel=1e-10)


def test_aitable_closed_form():
    rows = np.loadtxt(cases.os.path.join(cases.GOLDEN, "AiTable.dat"))
    x0 = cases.x0_closed_form(15, 0.9660499, 512.0, 256)
    np.testing.assert_allclose(rows[:, 1], x0[:, 0], rtol=1e-14)


@pytest.mark.parametrize("sphere", [0, 1])
def test_gaussian_invariants(sphere):
    N = 32
    cfg = cases.base_config(Nmesh=N, Nsample=24, Nmax=32, SphereMode=sphere)
    phi = orc.gaussian(cfg)
    inside, k, k2, _ = ngref.cut_mask(cfg)
    assert np.all(phi[~inside] == 0)
    assert np.all(phi[:, :, N // 2] == 0) and np.all(phi[N // 2] == 0) and np.all(phi[:, N // 2] == 0)
    # Hermitian k=0 plane, exactly (main.c:472-520)
    p0 = phi[:, :, 0]
    idx = (-np.arange(N)) % N
    np.testing.assert_array_equal(p0, np.conj(p0[idx][:, idx]))
    c = orc.cosmology(cfg)
    # |phi|^2 = Anorm k^ns (-ln u) (fac Beta / k^2)^2 with -ln u ~ Exp(1)
    e = (np.abs(phi[inside]) * k2[inside] / (c["fac"] * c["Beta"])) ** 2 / (c["Anorm"] * k[inside] ** cfg["PrimordialIndex"])
    assert abs(e.mean() - 1) < 0.05


@pytest.mark.parametrize("table", ["local", "equilateral", "shipped"])
def test_ng_potential_vs_regrouped_numpy(table):
    N = 16
    if table == "shipped":
        Nm = 15
        alpha = orc.alpha_table(Nm, cases.alpha_rows_file())
        x0 = cases.x0_file(Nm)
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=Nm, PrimordialIndex=0.9660499)
    else:
        Nm = 4
        rows = cases.local_rows() if table == "local" else cases.equilateral_rows()
        alpha = orc.alpha_table(Nm, rows)
        x0 = np.zeros((Nm, 2))
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=Nm)
    phig = orc.gaussian(cfg)
    c = orc.cosmology(cfg)
    import ctypes as C
    p = orc.make_params(cfg)
    cc = orc.OrcCosmo(**c)
    buf = np.ascontiguousarray(phig).view(np.float64).reshape(-1).copy()
    rc = orc.lib().orc_ng_potential(C.byref(p), C.byref(cc), np.ascontiguousarray(alpha.reshape(-1)),
                                    np.ascontiguousarray(x0.reshape(-1)), buf, None, None)
    assert rc == 0
    ours = buf.view(np.complex128).reshape(phig.shape)
    ref = ngref.potential(cfg, phig, alpha, x0)
    err = np.linalg.norm(ours - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err
    # the NG part is not negligible: Phi differs from phi_G
    assert np.linalg.norm(ours - phig) / np.linalg.norm(phig) > 1e-6


def test_sigma2_with_wdm_suppression_vs_quad():
    """WDM_On 1 multiplies P(k) by T_WDM(k)^2, T_WDM = (1 + (a k)^2.4)^(-5/1.2) with
    a = 0.048 ((Om - Ob)/0.4)^0.15 (h/0.65)^1.3 (1/m_keV)^1.15 (power.c:267-275)."""
    cfg = cases.base_config(WhichTransfer=0, WDM_On=1, WDM_PartMass_in_kev=0.5)
    R = 8.0
    ns = cfg["PrimordialIndex"]
    a = 0.048 * ((cfg["Omega"] - cfg["OmegaBaryon"]) / 0.4) ** 0.15 * (cfg["HubbleParam"] / 0.65) ** 1.3 \
        * (1 / cfg["WDM_PartMass_in_kev"]) ** 1.15

    def integrand(k):
        kr = R * k
        w = 3 * (np.sin(kr) / kr ** 3 - np.cos(kr) / kr ** 2)
        return 4 * np.pi * k * k * w * w * k ** ns * (1 + (a * k) ** 2.4) ** (-2 * 5 / 1.2)

    ref = integrate.quad(integrand, 1e-12, 500.0 / R, limit=2000, epsrel=1e-11)[0]
    ours = orc.lib().orc_sigma2_tophat(orc.make_params(cfg), 1.0, R)
    assert ours == pytest.approx(ref, rel=1e-7)
    plain = orc.lib().orc_sigma2_tophat(orc.make_params(dict(cfg, WDM_On=0)), 1.0, R)
    assert ours < 0.9 * plain  # a 0.5 keV particle suppresses sigma_8 visibly


def test_eisenstein_hu_transfer_limits_and_scalings():
    """The no-wiggle Eisenstein & Hu (1998, ApJ 496, 605) transfer function of power.c:519-548,
    pinned by what its published form fixes rather than by its formula:
    * no baryons (OmegaBaryon 1e-14; 0 would select power.c's 0.04 h^2 default): T depends on k
      only through q = k Theta_2.7^2 / (Omega h) (eq. 28 with Gamma = Omega h), so two
      cosmologies of equal Omega h agree and others follow by rescaling k;
    * q -> 0: (1 - T) / q^2 -> (14.2 + 731) / ln(2e); q -> oo: T 14.2 q^2 / ln(1.8 q) -> 1 (eq. 29);
    * baryons rescale Gamma by alpha_Gamma (eq. 31) above the sound horizon s (eq. 26) and not
      below it: T_b(k) = T_0(k / a_eff) with a_eff = alpha + (1 - alpha) / (1 + (0.43 k s)^4) (eq. 30).
    """
    theta2 = (2.728 / 2.7) ** 2

    def T(k, **kw):
        return orc.transfer(cases.base_config(WhichTransfer=1, **kw), k)

    ks = np.logspace(-4, 2, 25)
    cdm = dict(OmegaBaryon=1e-14)
    a = [T(k, Omega=0.3, HubbleParam=0.7, **cdm) for k in ks]
    b = [T(k, Omega=0.25, HubbleParam=0.84, **cdm) for k in ks]  # same Omega h
    np.testing.assert_allclose(a, b, rtol=1e-12)
    # Omega h = 0.18: k scaled by 0.18 / 0.21 gives the same q
    c = [T(k * 0.18 / 0.21, Omega=0.3, HubbleParam=0.6, **cdm) for k in ks]
    np.testing.assert_allclose(a, c, rtol=1e-12)

    gam = 0.3 * 0.7
    for q in [1e-5, 3e-6]:
        t = T(q * gam / theta2, Omega=0.3, HubbleParam=0.7, **cdm)
        assert (1 - t) / q ** 2 == pytest.approx(745.2 / np.log(2 * np.e), rel=2e-3)
    for q in [1e5, 1e6]:
        t = T(q * gam / theta2, Omega=0.3, HubbleParam=0.7, **cdm)
        assert t * 14.2 * q * q / np.log(1.8 * q) == pytest.approx(1.0, rel=1e-4)

    om, h, ob = 0.3, 0.7, 0.05
    omh2, obh2 = om * h * h, ob * h * h
    fb = obh2 / omh2
    alpha = 1 - 0.328 * np.log(431 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb * fb
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1 + 10 * obh2 ** 0.75) * h  # Mpc/h
    for x in [1e-2, 1.0, 30.0]:  # 0.43 k s
        k = x / (0.43 * s)
        aeff = alpha + (1 - alpha) / (1 + x ** 4)
        tb = T(k, Omega=om, HubbleParam=h, OmegaBaryon=ob)
        t0 = T(k / aeff, Omega=om, HubbleParam=h, **cdm)
        assert tb == pytest.approx(t0, rel=1e-9)
    assert T(1e-7, Omega=om, HubbleParam=h, OmegaBaryon=ob) == pytest.approx(1.0, abs=1e-9)
