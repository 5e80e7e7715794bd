"""Bispectrum templates as shape functions on the tetrapyd grid -- input
generation for tests and the benchmark (the decomposition's shapes).

S(k1, k2, k3) = (k1 k2 k3)^2 B(k1, k2, k3) with B the sum of the raw template
over the six orderings of (k1, k2, k3) (Aarambam/basis/decompose.py:219-271).
"""
import itertools

import numpy as np
from scipy import special


def scalar_i(k1, k2, k3, n_s, mass):
    """Scalar-I collider template of Sohn et al. 2024, eq. 2.15 (the dot(phi)^2 sigma
    interaction), with each k rescaled to k^((4 - n_s)/3) and the normalisation at the
    equilateral point, as Aarambam/models/Colliders.py:23-72 evaluates it."""
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
        # x3 is the soft leg, x1 + x2 the hard pair; returns the unnormalised shape
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


def shape(grid, raw):
    """(k1 k2 k3)^2 times the raw template summed over the orderings, at the tetrapyd points."""
    k = grid["k"]
    ks = [k[grid["i1"]], k[grid["i2"]], k[grid["i3"]]]
    B = 0
    for p in itertools.permutations(range(3)):
        B = B + raw(ks[p[0]], ks[p[1]], ks[p[2]])
    return (B * (ks[0] * ks[1] * ks[2]) ** 2)[None, :]


def decompose_bad_modes(grid):
    """The modes decompose.py:125-129 drops besides the ones basis_expansion always drops."""
    p1, p2, p3 = grid["p1"], grid["p2"], grid["p3"]
    return (p1 < 4) & ~(((p1 == 1) & (p2 == 1) & (p3 == 1)) | ((p1 == 2) & (p2 == 1) & (p3 == 0))
                        | ((p1 == 3) & (p2 == 0) & (p3 == 0)))
