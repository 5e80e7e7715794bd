"""The modal decomposition that makes the IC stage's coefficient tables, with
its contractions and its basis expansion on the device.

Mirrors the reference's decomposition path (Aarambam/basis/decompose.py with
Aarambam/basis/cmbbest.pyx): the Basis grid (mode functions, 1-D k grid and
weights, tetrapyd points, 3-D mode indices), the shape function of a
template, the basis expansion with the reference's settings, and the
AlphaTable / AiTable files.  The template itself (the shape function) is the
caller's, as in the reference, where models supply raw_bispectrum.
"""
import itertools
import os

import numpy as np
from scipy import special


def basis_grid(N_modes, n_s, k_min, k_max, k_grid_size=150):
    """The grid of cmbbest.Basis (cmbbest.pyx:57-98): dict of w, i1, i2, i3 (tetrapyd points),
    p1, p2, p3 (3-D modes, p1 >= p2 >= p3), evals (N_modes x k_grid_size), k."""
    scale = (4 - n_s) / 3
    # uniform grid and its weights (cmbbest.pyx:146-167)
    k = np.linspace(k_min, k_max, k_grid_size)
    k_median = np.sqrt(k_min * k_max)
    kw = np.where(k > k_median * 2, 1 / np.sqrt(k), 1 / k)
    kw[k < 2 * k_min] = 1 / np.sqrt(k[k < 2 * k_min])
    # tetrapyd points k1 >= k2 >= k3 with k2 + k3 >= k1 (cmbbest.pyx:170-214)
    ar = np.arange(k_grid_size)
    A, B, Cc = (g.reshape(-1) for g in np.meshgrid(ar, ar, ar, indexing="ij"))  # a major, c minor
    t = np.stack([A, B, Cc])[:, (A >= B) & (B >= Cc)]
    keep = k[t[1]] + k[t[2]] >= k[t[0]]
    i1, i2, i3 = t[:, keep]
    w = np.prod(kw[np.stack([i1, i2, i3])], axis=0)
    w = w / np.sum(w)  # cmbbest.pyx:74-83
    # mode functions (cmbbest.pyx:103-144): monomials for p <= 3, shifted Legendre polynomials above
    kmin_s, kmax_s = np.power(k_min, scale), np.power(k_max, scale)
    kbar = -1 + 2 / (kmax_s - kmin_s) * (np.power(k, scale) - kmin_s)
    kbar_zp = -1 - 2 * kmin_s / (kmax_s - kmin_s)
    evals = np.zeros((N_modes, k_grid_size))
    for p in range(N_modes):
        if p <= 3:
            evals[p] = np.power(k, 2 + (4 - n_s) / 3 * (p - 3))
        else:
            evals[p] = special.legendre(p - 1)(kbar) - special.eval_legendre(p - 1, kbar_zp)
    # 3-D mode indices (cmbbest.pyx:216-235)
    pi = np.arange(N_modes)
    P1, P2, P3 = np.meshgrid(pi, pi, pi, indexing="ij")
    o = (P1 >= P2) & (P2 >= P3)
    return dict(w=w, i1=i1, i2=i2, i3=i3, p1=P1[o], p2=P2[o], p3=P3[o], evals=evals, k=k)


def shape_values(grid, raw_bispectrum):
    """S = (k1 k2 k3)^2 sum over the six orderings of raw_bispectrum, at the tetrapyd points
    (decompose.py:219-271)."""
    k = grid["k"]
    ks = [k[grid["i1"]], k[grid["i2"]], k[grid["i3"]]]
    B = 0
    for a, b, c in itertools.permutations(range(3)):
        B = B + raw_bispectrum(ks[a], ks[b], ks[c])
    return B * (ks[0] * ks[1] * ks[2]) ** 2


def dropped_modes(grid):
    """The low-order modes decompose.py:125-129 leaves out of the expansion."""
    p1, p2, p3 = grid["p1"], grid["p2"], grid["p3"]
    return (p1 < 4) & ~(((p1 == 1) & (p2 == 1) & (p3 == 1)) | ((p1 == 2) & (p2 == 1) & (p3 == 0))
                        | ((p1 == 3) & (p2 == 0) & (p3 == 0)))


def ai_table(N_modes, n_s, k_min, k_max):
    """AiTable rows (n, x0, x1): x0 = P_n at the zero point of the k map, x1 = 0 (decompose.py:176-216)."""
    scale = (4 - n_s) / 3
    kmin_s, kmax_s = np.power(k_min, scale), np.power(k_max, scale)
    kbar_zp = -1 - 2 * kmin_s / (kmax_s - kmin_s)
    x0 = np.array([special.eval_legendre(p, kbar_zp) for p in range(N_modes)])
    return np.stack([np.arange(N_modes), x0, np.zeros(N_modes)], axis=1)


def decompose(ctx, raw_bispectrum, N_modes, n_s, Lbox, Nmax, ModeTol, MaxModeCount, k_grid_size=150):
    """BasicBasisDecompose.generate_alpha + generate_Ai (decompose.py:84-216) with the expansion on
    the device: returns dict(AlphaTable, AiTable, correlation, mse, selected)."""
    k_min, k_max = 2 * np.pi / Lbox, 2 * np.pi / Lbox * Nmax
    g = basis_grid(N_modes, n_s, k_min, k_max, k_grid_size)
    S = shape_values(g, raw_bispectrum)[None, :]
    alpha, _, corr, mse, sel = ctx.mode_basis_expansion(g, S, rtol=1e-10, Niter=1e3, bad_modes=dropped_modes(g),
                                                        RR_tol=ModeTol, max_modes=MaxModeCount)
    table = np.stack([g["p1"], g["p2"], g["p3"], alpha[0]], axis=1)
    return dict(AlphaTable=table, AiTable=ai_table(N_modes, n_s, k_min, k_max), correlation=float(corr[0]),
                mse=float(mse[0]), selected=sel[0], grid=g)


def write_tables(result, directory):
    """AlphaTable.dat (p1 p2 p3 alpha) and AiTable.dat (n x0 x1) as decompose.py:297-321 writes them."""
    np.savetxt(os.path.join(directory, "AlphaTable.dat"), result["AlphaTable"], fmt="%d %d %d %.17g")
    np.savetxt(os.path.join(directory, "AiTable.dat"), result["AiTable"], fmt="%d %.17g %.17g")


__all__ = ["basis_grid", "shape_values", "dropped_modes", "ai_table", "decompose", "write_tables"]
