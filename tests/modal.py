"""Inputs of the modal decomposition contractions, built the way the reference's
Basis class builds them (Aarambam/basis/cmbbest.pyx): Legendre mode functions
with the four monomial modes (103-144), the uniform k grid and its weights
(146-167), the tetrapyd points k1 >= k2 >= k3, k2 + k3 >= k1 (169-214), the
product weights normalised to unit sum (74-83) and the ordered 3-D mode indices
p1 >= p2 >= p3 (216-235).  Test inputs only.
"""
import numpy as np
from scipy.special import eval_legendre


def mode_functions(k, p_max, k_min, k_max, n_s):
    scale = (4 - n_s) / 3
    fact = 2 / (k_max ** scale - k_min ** scale)
    kbar = -1 + fact * (k ** scale - k_min ** scale)
    kbar_zp = -1 - 2 * k_min ** scale / (k_max ** scale - k_min ** scale)
    ev = np.zeros((p_max, len(k)))
    for p in range(p_max):
        if p <= 3:
            ev[p] = k ** (2 + (4 - n_s) / 3 * (p - 3))
        else:
            ev[p] = eval_legendre(p - 1, kbar) - eval_legendre(p - 1, kbar_zp)
    return ev


def basis_grid(p_max=8, k_grid_size=40, k_min=2 * np.pi / 1000, k_max=2 * np.pi / 1000 * 256, n_s=0.9660499):
    k = np.linspace(k_min, k_max, k_grid_size)
    k_median = np.sqrt(k_min * k_max)
    kw = np.where(k > k_median * 2, 1 / np.sqrt(k), 1 / k)
    kw[k < 2 * k_min] = 1 / np.sqrt(k[k < 2 * k_min])
    t = np.array([[a, b, c] for a in range(k_grid_size) for b in range(a + 1) for c in range(b + 1)]).T
    keep = k[t[1]] + k[t[2]] >= k[t[0]]
    i1, i2, i3 = t[:, keep]
    w = kw[i1] * kw[i2] * kw[i3]
    w = w / w.sum()
    pi = np.arange(p_max)
    P1, P2, P3 = np.meshgrid(pi, pi, pi, indexing="ij")
    o = (P1 >= P2) & (P2 >= P3)
    return dict(w=w, i1=i1, i2=i2, i3=i3, p1=P1[o], p2=P2[o], p3=P3[o],
                evals=mode_functions(k, p_max, k_min, k_max, n_s), k=k)
