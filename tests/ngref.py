"""Independent numpy restatement of the modal NG sum, regrouped by outer index.

Used to pin the oracle's pairwise loops (Aarambam/LPT/Basis/main.c:564-1151):
mode fields f_n = q_n(k) phi(k) with q_n = g_n(k) k^(4-ns), weights
g_n = k^(s(n-3)) (n<=3), (P_{n-1}(xbar) - x0[n-1])/k^2 (n>=4), s=(4-ns)/3;
NG = sum_{i<=j<=K} alpha_ijK g_K FFT[f_i f_j] + Beta terms; the cut and the
origin are excluded.  FFT conventions are numpy's, which equal FFTW's
real-transform conventions (imaginary parts of the z bins 0 and N/2 ignored
by the inverse).
"""
import numpy as np
from scipy import special


def kgrid(N, Box):
    idx = np.arange(N)
    kv = np.where(idx < N // 2, idx * 2 * np.pi / Box, -(N - idx) * 2 * np.pi / Box)
    kz = kv[: N // 2 + 1].copy()
    kz[N // 2] = -(N - N // 2) * 2 * np.pi / Box
    return kv, kz


def cut_mask(cfg):
    N, Box, Ns = cfg["Nmesh"], cfg["Box"], cfg["Nsample"]
    kv, kz = kgrid(N, Box)
    KX, KY, KZ = np.meshgrid(kv, kv, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k = np.sqrt(k2)
    if cfg.get("SphereMode", 0) == 1:
        inside = ~(k * Box / (2 * np.pi) > Ns // 2)
    else:
        inside = ~((np.abs(KX) * Box / (2 * np.pi) > Ns // 2) | (np.abs(KY) * Box / (2 * np.pi) > Ns // 2)
                   | (np.abs(KZ) * Box / (2 * np.pi) > Ns // 2))
    inside[0, 0, 0] = False
    return inside, k, k2, (KX, KY, KZ)


def weights(cfg, x0, k, k2, n):
    ns = cfg["PrimordialIndex"]
    s = (4 - ns) / 3
    with np.errstate(divide="ignore", invalid="ignore"):
        if n <= 3:
            return k ** (s * (n - 3))
        kmin = 2 * np.pi / cfg["Box"]
        kmax = kmin * cfg["Nmax"]
        xb = -1 + 2 * (k ** s - kmin ** s) / (kmax ** s - kmin ** s)
        return (special.eval_legendre(n - 1, xb) - x0[n - 1, 0]) / k2


def ng_terms(alpha, Nm):
    """{K: [(i, j, coef)]} regrouped by the outer index, Beta terms folded in."""
    terms = {}
    for i in range(Nm):
        for j in range(i, Nm):
            for K in range(j, Nm):
                a = alpha[i, j, K]
                if a != 0:
                    terms.setdefault(K, []).append((i, j, a))
    a012, a111 = alpha[0, 1, 2], alpha[1, 1, 1]
    t012 = 0.0 if a012 == 0 else -a111 / a012
    if t012 != 0:
        terms.setdefault(2, []).append((0, 1, -a012 * t012))
        terms.setdefault(1, []).append((0, 2, a012 * t012))
    return terms


def potential(cfg, phig, alpha, x0):
    N = cfg["Nmesh"]
    Nm = cfg["N_modes"]
    ns = cfg["PrimordialIndex"]
    inside, k, k2, _ = cut_mask(cfg)
    terms = ng_terms(alpha, Nm)
    modes = sorted({t[0] for v in terms.values() for t in v} | {t[1] for v in terms.values() for t in v})
    fields = {}
    for n in modes:
        q = np.where(inside, weights(cfg, x0, k, k2, n) * k ** (4 - ns), 0.0)
        fields[n] = np.fft.irfftn(q * phig, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    ng = np.zeros_like(phig)
    for K, lst in terms.items():
        h = np.zeros((N, N, N))
        for i, j, c in lst:
            h += c * fields[i] * fields[j]
        g = np.where(inside, weights(cfg, x0, k, k2, K), 0.0)
        ng += g * np.fft.rfftn(h)
    phi = phig + cfg["Fnl"] * ng / N ** 3
    phi[0, 0, 0] = 0
    # the reference's two C2R/R2C round trips project the z-bin 0 and N/2 planes
    return np.fft.rfftn(np.fft.irfftn(phi, s=(N, N, N), axes=(0, 1, 2)))
