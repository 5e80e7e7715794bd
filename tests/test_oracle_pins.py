"""Pins the oracle's LPT half, particle stage, glass tiling, Gaussian draw mapping and
seed-table order against closed forms and hand enumerations (CPU only).

What each check fixes, independently of the oracle's own formulas:
* seed table (main.c:317-349): the cell filled by every draw, enumerated by hand for N = 4, 6;
* Gaussian potential (main.c:382-524): each mode's phase is 2 pi u_2k (+ pi with PhaseFlip) and
  its power -ln u_2k+1 times the spectrum, with u the (i, j) pencil's own stream (pinned by GSL's
  known answers), draws consumed for cut and Nyquist modes;
* ZA and 2LPT fields (main.c:1158-1381): plane waves whose displacements, 2LPT source and
  CIC deconvolution have closed forms;
* CIC readout (main.c:1440-1516): nodes read back exactly, linear fields interpolate exactly,
  the clamps at i, j, k == N and the periodic wrap;
* glass tiling (read_glass.c:176-214): positions and IDs of a 2 x 2 x 2 tiling by hand;
* WDM thermal speeds (power.c:624-707): drand48 against the C library's, the speeds' moments
  against the Fermi-Dirac closed forms, isotropic directions.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases

# (row, column) filled by each draw of the seed table, in draw order, written out by hand
# from the eight loops of main.c:321-349 for shells i = 0 .. N/2 - 1.
SEED_ORDER = {
    4: [(0, 0), (3, 0), (0, 3), (3, 3),
        (1, 0), (0, 1), (1, 1), (2, 0), (3, 1), (2, 1), (1, 3), (0, 2), (1, 2), (2, 3), (3, 2), (2, 2)],
    6: [(0, 0), (5, 0), (0, 5), (5, 5),
        (1, 0), (0, 1), (1, 1), (4, 0), (5, 1), (4, 1), (1, 5), (0, 4), (1, 4), (4, 5), (5, 4), (4, 4),
        (2, 0), (2, 1), (0, 2), (1, 2), (2, 2), (3, 0), (3, 1), (5, 2), (4, 2), (3, 2),
        (2, 5), (2, 4), (0, 3), (1, 3), (2, 3), (3, 5), (3, 4), (5, 3), (4, 3), (3, 3)],
}


@pytest.mark.parametrize("N", [4, 6])
def test_seedtable_full_fill_order(N):
    order = SEED_ORDER[N]
    assert len(order) == N * N and len(set(order)) == N * N
    for seed in (42, 987654):
        u = orc.ranlxd_draws(seed, N * N)
        st = orc.seedtable(N, seed)
        want = np.zeros((N, N), np.uint32)
        for t, (r, c) in enumerate(order):
            want[r, c] = np.uint32(0x7FFFFFFF * u[t])  # unsigned int = 0x7fffffff * uniform
        np.testing.assert_array_equal(st, want)


def _kvec(N, Box, idx):
    return np.where(idx < N // 2, idx, idx - N) * 2 * np.pi / Box


@pytest.mark.parametrize("kw", [dict(SphereMode=0, Nsample=10), dict(SphereMode=1, Nsample=12),
                                dict(SphereMode=0, Nsample=16, PhaseFlip=1),
                                dict(SphereMode=1, Nsample=10, FixedAmplitude=1)])
def test_gaussian_modes_follow_their_pencil_streams(kw):
    """Every mode against the draws of its own pencil: phase 2 pi u[2k] (+ pi), and
    |phi|^2 k^4 / (Anorm k^ns (fac Beta)^2) = -ln u[2k+1] (1 with FixedAmplitude).  The k = 0
    plane takes pencil (i, j) for i = 0, j < N/2 and 0 < i < N/2, and the conjugate of
    pencil (N - i, N - j) elsewhere; every other mode is zero.  With the cut at Nsample < N,
    pencils whose low k are kept and high k cut, and (sphere) pencils cut throughout, run
    before kept ones, so a draw skipped for a cut or Nyquist mode shifts what follows."""
    N, Box = 16, 1000.0
    cfg = cases.base_config(Nmesh=N, Nmax=N, Box=Box, Seed=2718, **kw)
    phi = orc.gaussian(cfg)
    c = orc.cosmology(cfg)
    st = orc.seedtable(N, cfg["Seed"])
    ns, half = cfg["PrimordialIndex"], N // 2
    flip = np.pi if cfg.get("PhaseFlip") else 0.0
    want = np.zeros_like(phi)
    kc = lambda i: (i if i < half else i - N) * 2 * np.pi / Box

    def inside(kv):
        if cfg["SphereMode"]:
            return np.sqrt(sum(x * x for x in kv)) * Box / (2 * np.pi) <= cfg["Nsample"] // 2
        return all(abs(x) * Box / (2 * np.pi) <= cfg["Nsample"] // 2 for x in kv)

    n_kept = n_plane = 0
    for i in range(N):
        for j in range(N):
            u = orc.ranlxd_draws(int(st[i, j]), N)
            assert np.all(u[1::2] > 0)  # no redraw needed for these streams
            for k in range(half):
                if i == half or j == half or (i == j == k == 0):
                    continue
                kv = (kc(i), kc(j), kc(k))
                if not inside(kv):
                    continue
                k2 = sum(x * x for x in kv)
                power = c["Anorm"] * np.sqrt(k2) ** ns * (1.0 if cfg["FixedAmplitude"] else -np.log(u[2 * k + 1]))
                val = np.sqrt(power) * c["fac"] * c["Beta"] / k2 * np.exp(1j * (2 * np.pi * u[2 * k] + flip))
                if k > 0:
                    want[i, j, k] = val
                    n_kept += 1
                elif i == 0 and j < half:
                    want[0, j, 0], want[0, (N - j) % N, 0] = val, np.conj(val)
                    n_plane += 1
                elif 0 < i < half:
                    want[i, j, 0], want[N - i, (N - j) % N, 0] = val, np.conj(val)
                    n_plane += 1
    assert n_kept > 100 and n_plane > 10
    np.testing.assert_array_equal(phi == 0, want == 0)
    nz = want != 0
    np.testing.assert_allclose(phi[nz], want[nz], rtol=1e-12, atol=0)


def _sinc(x):
    return np.sin(x) / x if x != 0 else 1.0


def _lpt_setup(N=16, Box=1000.0):
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, Box=Box)
    c = orc.cosmology(cfg)
    tw = lambda kmag: orc.transfer(cfg, kmag) / c["Dplus"] / c["Beta"]
    x = np.arange(N) * Box / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return cfg, tw, (X, Y, Z)


@pytest.mark.parametrize("theta", [0.0, 0.7, -2.1])
def test_lpt_single_axis_wave(theta):
    """Phi = A cos(k0 z + theta): psi = T(k0)/(D+ Beta) grad Phi times the CIC window
    sinc(k0 L / 2N)^-2, and the 2LPT source of a one-dimensional wave is exactly zero."""
    N, Box, m, A = 16, 1000.0, 3, 2.5e-3
    cfg, tw, (X, Y, Z) = _lpt_setup(N, Box)
    k0 = m * 2 * np.pi / Box
    spec = np.zeros((N, N, N // 2 + 1), np.complex128)
    spec[0, 0, m] = A / 2 * np.exp(1j * theta)
    d = orc.lpt_fields(cfg, spec)
    s = _sinc(k0 * Box / (2 * N)) ** -2
    want_z = -k0 * tw(k0) * A * np.sin(k0 * Z + theta) * s
    scale = np.abs(want_z).max()
    np.testing.assert_allclose(d[2], want_z, rtol=0, atol=1e-12 * scale)
    assert np.all(d[0] == 0) and np.all(d[1] == 0)
    for f in d[3:]:
        assert np.all(f == 0)


def test_lpt_wave_along_complex_axis():
    """The same along x, where the k = 0 plane carries the conjugate pair (m, 0, 0), (N - m, 0, 0)."""
    N, Box, m, A = 16, 1000.0, 5, 1e-3
    cfg, tw, (X, Y, Z) = _lpt_setup(N, Box)
    k0 = m * 2 * np.pi / Box
    spec = np.zeros((N, N, N // 2 + 1), np.complex128)
    spec[m, 0, 0] = A / 2 * np.exp(0.4j)
    spec[N - m, 0, 0] = A / 2 * np.exp(-0.4j)
    d = orc.lpt_fields(cfg, spec)
    want_x = -k0 * tw(k0) * A * np.sin(k0 * X + 0.4) * _sinc(k0 * Box / (2 * N)) ** -2
    np.testing.assert_allclose(d[0], want_x, rtol=0, atol=1e-12 * np.abs(want_x).max())
    assert np.all(d[1] == 0) and np.all(d[2] == 0)
    for f in d[3:]:
        assert np.all(f == 0)


def test_lpt_diagonal_wave_cic_window():
    """A wave along (2, 3, 1): each displacement component carries the product of the three
    axis windows sinc(k_a L / 2N)^-2, and the source cancels to rounding."""
    N, Box, A = 16, 1000.0, 1e-3
    cfg, tw, (X, Y, Z) = _lpt_setup(N, Box)
    m = np.array([2, 3, 1])
    kv = m * 2 * np.pi / Box
    kmag = np.sqrt((kv ** 2).sum())
    spec = np.zeros((N, N, N // 2 + 1), np.complex128)
    spec[2, 3, 1] = A / 2
    d = orc.lpt_fields(cfg, spec)
    window = np.prod([_sinc(k * Box / (2 * N)) ** -2 for k in kv])
    phase = kv[0] * X + kv[1] * Y + kv[2] * Z
    for a in range(3):
        want = -kv[a] * tw(kmag) * A * np.sin(phase) * window
        np.testing.assert_allclose(d[a], want, rtol=0, atol=1e-12 * np.abs(want).max())
    psi = np.abs(d[0]).max()
    for f in d[3:]:  # N^3 times a displacement of order k psi^2
        assert np.abs(f).max() < 1e-10 * N ** 3 * kmag * psi * psi


def test_lpt_two_orthogonal_waves_source_closed_form():
    """Phi = A cos(k1 x) + B cos(k2 y): phi_,xx phi_,yy is the only source term,
    S = A B k1^2 k2^2 t1 t2 cos(k1 x) cos(k2 y) with t = T/(D+ Beta), so
    psi2 = grad lap^-1 S = S0 / (k1^2 + k2^2) (k1 sin cos, k2 cos sin, 0), deconvolved by the
    window of (k1, k2, 0); the fields come back N^3 times larger (R2C is unnormalised)."""
    N, Box, m1, m2, A, B = 16, 1000.0, 2, 3, 1.3e-3, -0.8e-3
    cfg, tw, (X, Y, Z) = _lpt_setup(N, Box)
    k1, k2 = m1 * 2 * np.pi / Box, m2 * 2 * np.pi / Box
    spec = np.zeros((N, N, N // 2 + 1), np.complex128)
    spec[m1, 0, 0] = spec[N - m1, 0, 0] = A / 2
    spec[0, m2, 0] = spec[0, N - m2, 0] = B / 2
    d = orc.lpt_fields(cfg, spec)
    t1, t2 = tw(k1), tw(k2)
    s1, s2 = _sinc(k1 * Box / (2 * N)) ** -2, _sinc(k2 * Box / (2 * N)) ** -2
    np.testing.assert_allclose(d[0], -k1 * t1 * A * np.sin(k1 * X) * s1, rtol=0, atol=1e-12 * k1 * t1 * abs(A))
    np.testing.assert_allclose(d[1], -k2 * t2 * B * np.sin(k2 * Y) * s2, rtol=0, atol=1e-12 * k2 * t2 * abs(B))
    assert np.all(d[2] == 0)
    S0 = A * B * k1 ** 2 * k2 ** 2 * t1 * t2
    q = S0 / (k1 ** 2 + k2 ** 2) * s1 * s2 * N ** 3
    w2x = q * k1 * np.sin(k1 * X) * np.cos(k2 * Y)
    w2y = q * k2 * np.cos(k1 * X) * np.sin(k2 * Y)
    tol = 1e-11 * abs(q) * max(k1, k2)
    np.testing.assert_allclose(d[3], w2x, rtol=0, atol=tol)
    np.testing.assert_allclose(d[4], w2y, rtol=0, atol=tol)
    np.testing.assert_allclose(d[5], 0, atol=tol)
    assert np.abs(d[3]).max() > 100 * tol


def _particle_cfg(N=16, Box=32.0):
    return cases.base_config(Nmesh=N, Nsample=N, Nmax=N, Box=Box)


def test_particles_at_nodes_read_grid_values():
    """Box = 2N: node positions are exact binary fractions, so every particle sits on a node
    (weight 1 on corner (i, j, k), 0 elsewhere) and reads that node's values exactly:
    vel = dis vp1 - 3/7 (dis2 / N^3) vp2, pos += dis - 3/7 dis2 / N^3, wrapped."""
    N, Box = 16, 32.0
    cfg = _particle_cfg(N, Box)
    c = orc.cosmology(cfg)
    rng = np.random.default_rng(5)
    disp = [rng.standard_normal((N, N, N)) * 0.3 for _ in range(3)]
    disp2 = [rng.standard_normal((N, N, N)) * 0.2 * N ** 3 for _ in range(3)]
    idx = rng.integers(0, N, size=(500, 3))
    pos0 = (idx * (Box / N)).astype(np.float32)
    pos, vel = orc.particles(cfg, pos0, disp, disp2)
    n3 = np.float32(N ** 3)
    for a in range(3):
        dis = disp[a][idx[:, 0], idx[:, 1], idx[:, 2]]
        dis2 = disp2[a][idx[:, 0], idx[:, 1], idx[:, 2]] / float(n3)
        np.testing.assert_array_equal(vel[:, a], (dis * c["vel_prefac"] - 3. / 7. * dis2 * c["vel_prefac2"]).astype(np.float32))
        p = pos0[:, a].astype(np.float64) + dis - 3. / 7. * dis2
        p = np.float32(p).astype(np.float64)
        np.testing.assert_array_equal(pos[:, a], np.fmod(np.fmod(p, Box) + Box, Box).astype(np.float32))


def test_particles_interpolate_linear_fields_exactly():
    """Trilinear weights reproduce a linear function inside a cell: with disp_a = c_a . (i, j, k)
    + d_a on the nodes and disp2 = 0, vel = vp1 (c_a . (u, v, w) + d_a), u = Pos N / Box.  The
    three gradients differ, so a corner taken from the wrong axis or weight fails."""
    N, Box = 16, 1000.0
    cfg = _particle_cfg(N, Box)
    c = orc.cosmology(cfg)
    g = np.arange(N, dtype=np.float64)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    coef = [(0.31, -0.17, 0.05, 0.2), (-0.02, 0.44, -0.29, -0.1), (0.13, 0.07, 0.61, 0.0)]
    disp = [a * I + b * J + cc * K + d for a, b, cc, d in coef]
    zero = [np.zeros((N, N, N))] * 3
    rng = np.random.default_rng(9)
    pos0 = rng.uniform(0, (N - 1) * Box / N, size=(2000, 3)).astype(np.float32)
    U = pos0.astype(np.float64) / Box * N
    _, vel = orc.particles(cfg, pos0, disp, zero)
    for ax, (a, b, cc, d) in enumerate(coef):
        want = (a * U[:, 0] + b * U[:, 1] + cc * U[:, 2] + d) * c["vel_prefac"]
        np.testing.assert_allclose(vel[:, ax], want, rtol=2e-7, atol=2e-7 * np.abs(want).max())


def test_particles_clamp_at_upper_face_and_wrap():
    """A particle at exactly Box on an axis has index N, which the reference clamps to N - 1
    with weight 1 on the next plane, the periodic image (halo) plane 0; and a displacement
    across the box edge wraps into [0, Box)."""
    N, Box = 16, 32.0
    cfg = _particle_cfg(N, Box)
    c = orc.cosmology(cfg)
    rng = np.random.default_rng(3)
    disp = [rng.standard_normal((N, N, N)) for _ in range(3)]
    zero = [np.zeros((N, N, N))] * 3
    cases_ = [((Box, 4.0, 6.0), (0, 2, 3)), ((8.0, Box, 6.0), (4, 0, 3)), ((8.0, 4.0, Box), (4, 2, 0)),
              ((Box, Box, Box), (0, 0, 0))]
    pos0 = np.array([p for p, _ in cases_], np.float32)
    _, vel = orc.particles(cfg, pos0, disp, zero)
    for n, (_, (i, j, k)) in enumerate(cases_):
        for a in range(3):
            assert vel[n, a] == np.float32(disp[a][i, j, k] * c["vel_prefac"])
    # wrap: node particles near the faces pushed across them
    flat = [np.zeros((N, N, N)) for _ in range(3)]
    flat[0][N - 1, :, :] = 3.0   # x = 30 -> 33 -> 1
    flat[1][0, 0, :] = -1.5      # y = 0 -> -1.5 -> 30.5
    pos0 = np.array([[30.0, 6.0, 6.0], [0.0, 0.0, 8.0]], np.float32)
    pos, _ = orc.particles(cfg, pos0, flat, zero)
    assert pos[0, 0] == np.float32(1.0) and pos[1, 1] == np.float32(30.5)
    assert np.all((pos >= 0) & (pos < Box))


def test_glass_tiling_order_and_ids_by_hand():
    """read_glass.c:176-214 with GlassTileFac 2: tiles in i, j, k order (k fastest), the glass in
    file order inside each tile, x = g / BoxSize_g (Box / T) + tile (Box / T), IDs from 1."""
    cfg = cases.base_config(Nmesh=4, GlassTileFac=2, Box=20.0)
    glass = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], np.float32)
    pos, ids = orc.glass_tile(cfg, glass, 10.0, with_ids=True)
    want = [(1, 2, 3), (4, 5, 6), (1, 2, 13), (4, 5, 16), (1, 12, 3), (4, 15, 6), (1, 12, 13), (4, 15, 16),
            (11, 2, 3), (14, 5, 6), (11, 2, 13), (14, 5, 16), (11, 12, 3), (14, 15, 6), (11, 12, 13),
            (14, 15, 16)]
    np.testing.assert_array_equal(pos, np.array(want, np.float32))
    np.testing.assert_array_equal(ids, np.arange(1, 17))


def test_drand48_matches_libc():
    """The oracle's drand48 (POSIX's 48-bit generator, which the reference's thermal speeds draw
    from without seeding) against the C library's own drand48 reset to the default state."""
    import ctypes
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    libc.seed48.argtypes = [ctypes.POINTER(ctypes.c_ushort)]
    libc.seed48((ctypes.c_ushort * 3)(0x330E, 0xABCD, 0x1234))  # X0 = 0x1234ABCD330E, POSIX's default
    want = np.array([libc.drand48() for _ in range(5000)])
    np.testing.assert_array_equal(orc.drand48(5000), want)


def test_wdm_thermal_speeds_follow_fermi_dirac():
    """Kicks added to zero velocities: speeds |v| / V0 distributed as x^2 / (e^x + 1) on [0, 20],
    whose moments are (1 - 2^-n) n! zeta(n + 1) / (1.5 zeta(3)): mean 3.1514, rms 3.5971 (the
    constant power.c:653 prints); directions isotropic; V0 = 0.012 (1 + z) ((Om - Ob)/0.3)^(1/3)
    (h/0.65)^(2/3) m^(-4/3) km/s, times sqrt(1 + z) in Gadget units (power.c:649-661)."""
    from scipy.special import zeta
    cfg = cases.base_config(WDM_On=1, WDM_PartMass_in_kev=0.7, Redshift=40.0)
    n = 200000
    v = orc.wdm_thermal(cfg, np.zeros((n, 3), np.float32)).astype(np.float64)
    z = cfg["Redshift"]
    v0 = (0.012 * (1 + z) * ((cfg["Omega"] - cfg["OmegaBaryon"]) / 0.3) ** (1 / 3) * (cfg["HubbleParam"] / 0.65) ** (2 / 3)
          * (1 / cfg["WDM_PartMass_in_kev"]) ** (4 / 3) * 1e5 / cfg["UnitVelocity_in_cm_per_s"] * np.sqrt(1 + z))
    s = np.linalg.norm(v, axis=1) / v0
    import math
    mom = lambda k: (1 - 2.0 ** -k) * math.factorial(k) * zeta(k + 1)
    mean, rms = mom(3) / mom(2), np.sqrt(mom(4) / mom(2))
    assert abs(rms - 3.59714) < 1e-4
    assert s.mean() == pytest.approx(mean, abs=0.02)
    assert np.sqrt((s * s).mean()) == pytest.approx(rms, abs=0.02)
    assert s.max() <= 20.0 + 1e-4
    u = v / np.linalg.norm(v, axis=1)[:, None]
    assert np.abs(u.mean(axis=0)).max() < 0.01
    np.testing.assert_allclose((u * u).mean(axis=0), 1 / 3, atol=0.01)
