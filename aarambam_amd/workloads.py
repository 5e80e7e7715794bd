"""Inputs of IC jobs built in memory: parameter sets, lattice glasses, closed-form modal tables
and the ScalarI template, for the benchmark and for callers without parameter files.

* `config` holds the reference parameter file's tags (Aarambam/LPT/Basis/read_param.c:27-184)
  with the defaults of a plain FP64 run (64^3, L = 1000 Mpc/h, Eisenstein-Hu P(k), Om 0.3,
  Ob 0.05, h 0.7, n_s 0.9649, sigma8 0.8, z 50, fNL 100, seed 1234).
* `lattice_glass` is the layout of Aarambam/defaults/dummy_glass_dmonly_64.dat: an n^3 lattice,
  z fastest, spacing box / n.
* The local and equilateral tables are the single separable triples the templates reduce to
  (Aarambam/models/SingleField.py:19-27, 45-78); X0 is the closed form of decompose.py:176-216.
* `scalar_i` is the collider template of Aarambam/models/Colliders.py:23-72 (the dot(phi)^2 sigma
  interaction of Sohn et al. 2024, eq. 2.15) as a raw bispectrum for aarambam_amd.decompose.
"""
import itertools

import numpy as np
from scipy import special

UNITS = dict(UnitLength_in_cm=3.085678e24, UnitMass_in_g=1.989e43, UnitVelocity_in_cm_per_s=1e5,
             InputSpectrum_UnitLength_in_cm=3.085678e24)


def config(**kw):
    cfg = dict(Nmesh=64, Nsample=64, Nmax=64, N_modes=8, Seed=1234, SphereMode=0, FixedAmplitude=0, PhaseFlip=0,
               WhichSpectrum=0, WhichTransfer=1, GlassTileFac=1, WDM_On=0, Box=1000.0, Omega=0.3,
               OmegaLambda=0.7, OmegaBaryon=0.05, HubbleParam=0.7, Sigma8=0.8, PrimordialIndex=0.9649,
               Redshift=50.0, Fnl=100.0, ShapeGamma=0.0, WDM_PartMass_in_kev=10.0, **UNITS)
    cfg.update(kw)
    return cfg


def lattice_glass(n=64, box=100.0):
    g = np.arange(n, dtype=np.float64) * (box / n)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)


def local_rows():
    """The local template: the triple (3, 0, 0) with alpha 1."""
    return np.array([[3, 0, 0, 1.0]])


def equilateral_rows():
    """The equilateral template: alpha(3,0,0) = -3, alpha(1,1,1) = -2, alpha(2,1,0) = 6."""
    return np.array([[3, 0, 0, -3.0], [1, 1, 1, -2.0], [2, 1, 0, 6.0]])


def le_rows():
    """Local + equilateral: the sum of the two tables (the (3, 0, 0) entries add to -2)."""
    merged = {}
    for n1, n2, n3, a in list(local_rows()) + list(equilateral_rows()):
        key = tuple(sorted((int(n1), int(n2), int(n3)), reverse=True))
        merged[key] = merged.get(key, 0.0) + float(a)
    return np.array([[*k, v] for k, v in sorted(merged.items())])


def dense12_rows(n_triples=60, seed=12):
    """A synthetic dense 12-mode table: 60 distinct triples p1 >= p2 >= p3 < 12 with
    alpha ~ N(0, 1) from PCG64(seed) (39 distinct product pairs)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    triples = [(a, b, c) for a in range(12) for b in range(a + 1) for c in range(b + 1)]
    pick = rng.choice(len(triples), size=n_triples, replace=False)
    return np.array([[*triples[i], rng.standard_normal()] for i in sorted(pick)])


def alpha_table(nm, rows):
    """N_modes^3 table with every row written to all six orderings (main.c:231-241)."""
    table = np.zeros((nm, nm, nm))
    for n1, n2, n3, a in np.asarray(rows, dtype=np.float64).reshape(-1, 4):
        for q in itertools.permutations((int(n1), int(n2), int(n3))):
            table[q] = a
    return table


def x0_table(nm, n_s, box, nmax):
    """X0 rows (x0, x1): x0[p] = P_p(-1 - 2 kmin^s / (kmax^s - kmin^s)), s = (4 - n_s) / 3,
    kmin = 2 pi / box, kmax = nmax kmin; x1 = 0."""
    s = (4 - n_s) / 3
    kmin = 2 * np.pi / box
    kmax = kmin * nmax
    kbar = -1 - 2 * kmin ** s / (kmax ** s - kmin ** s)
    x0 = np.zeros((nm, 2))
    for p in range(nm):
        x0[p, 0] = special.eval_legendre(p, kbar)
    return x0


def scalar_i(k1, k2, k3, n_s, mass):
    """The ScalarI raw bispectrum at (k1, k2, k3), each k rescaled to k^((4 - n_s)/3), normalised at
    the equilateral point."""
    s = (4 - n_s) / 3
    a, b, c = np.power(k1, s), np.power(k2, s), np.power(k3, s)
    mu = mass
    beta = mu * mu + 0.25
    ph = np.angle(special.gamma(2.5 + 1j * mu) * special.gamma(-1j * mu) * (1 + 1j * np.sinh(np.pi * mu)))
    amp3 = np.sqrt(np.pi ** 3 * beta * (beta + 2) / mu / np.sinh(2 * np.pi * mu))
    ch = np.cosh(np.pi * mu)
    q4 = (1 + 16 * mu * mu) / (1 + 8 * mu * mu)
    c24 = 2 * (3 + 68 * mu ** 2 + 384 * mu ** 4) / (1 + 8 * mu * mu) ** 2

    def terms(x1, x2, x3):
        # x3 is the soft leg, x1 + x2 the hard pair
        t = x1 + x2 + x3
        h = x1 + x2
        lg = np.log(t / h)
        first = (2 * x1 * x2 * x3 / (beta * t * t) / h
                 * (1 + 4 * x3 / (beta + 2) / h + (beta + 4) * x3 * x3 / (beta + 2) ** 2 / (h * h))
                 * np.power(t / h, -beta / (beta + 2)))
        second = (x1 * x2 * x3 / (6 * ch * h ** 3)
                  * (-2 * (2 * mu ** 4 - 1) * (x3 * x3 / (t * t) + x3 * (4 * t - x3) / (t * t) * lg + lg * lg)
                     + mu * mu * np.power(x3 / h, q4)
                     * (x3 * (6 * t - x3 + 8 * mu * mu * (8 * t - x3)) / (1 + 8 * mu * mu) / (t * t) + c24 * lg)))
        third = x1 * x2 / (h * h) * amp3 * np.sqrt(x3 / h) * np.cos(mu * np.log(x3 / 2 / h) + ph)
        return first + second + third

    one = np.ones(1)
    norm = 3 * terms(one, one, one)[0]
    return terms(a, b, c) / norm * 0.5 / (a * b * c) ** 2


__all__ = ["config", "lattice_glass", "local_rows", "equilateral_rows", "le_rows", "dense12_rows", "alpha_table",
           "x0_table", "scalar_i"]
