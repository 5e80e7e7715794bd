"""Bispectrum templates as shape functions on the tetrapyd grid -- input
generation for tests and the benchmark (the decomposition's shapes).

S(k1, k2, k3) = (k1 k2 k3)^2 B(k1, k2, k3) with B the sum of the raw template
over the six orderings of (k1, k2, k3) (Aarambam/basis/decompose.py:219-271).
"""
import itertools

from aarambam_amd.workloads import scalar_i  # noqa: F401  (Colliders.py:23-72)


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
