"""CPU restatement of the modal basis expansion -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and bench.py's CPU leg, never by the product package.
Follows Basis.basis_expansion of Aarambam/basis/cmbbest.pyx:320-490 with the
same library calls (scipy.sparse.linalg.cg, numpy.linalg.cond) and the mode
bispectra of Basis.evaluate_mode_bispectra (cmbbest.pyx:237-258).  The
covariance QQ comes from the C oracle's restatement of
compute_mode_bispectra_covariance (oracle.c), as the reference takes it from
tetrapyd.c.  The reference's joblib pool over candidates is a plain loop here.

Pinned in tests/test_modal.py by closed forms: a shape in the span of the
starting modes is reproduced exactly without a selection step, a shape in the
span of all modes with max_modes -1 is reproduced exactly, and a shape made of
one mode outside the starting set is recovered by the selection loop.  Which
mode the loop adds at each step (the ten most correlated, the smallest rise of
the condition number) has no closed form; it is pinned by vectors from the
reference's own cmbbest (built from its sources into oracle/_ref/ and run by
tests/golden/make_expansion_vectors.py): for ScalarI (132 modes picked) and the
equilateral template on a 40-point k grid with 16 modes, this restatement
returns the reference's coefficients bit for bit.
"""
import numpy as np
from scipy.sparse.linalg import cg as conjugate_gradient

from . import pyoracle


def mode_bispectra(grid):
    """Q[n, t], cmbbest.pyx:237-258."""
    ev = np.asarray(grid["evals"])
    p1, p2, p3 = grid["p1"], grid["p2"], grid["p3"]
    t1, t2, t3 = grid["i1"], grid["i2"], grid["i3"]
    f1, f2, f3 = ev[p1, :], ev[p2, :], ev[p3, :]
    s = (f1[:, t1] * f2[:, t2] * f3[:, t3] + f1[:, t1] * f2[:, t3] * f3[:, t2]
         + f1[:, t2] * f2[:, t1] * f3[:, t3] + f1[:, t2] * f2[:, t3] * f3[:, t1]
         + f1[:, t3] * f2[:, t1] * f3[:, t2] + f1[:, t3] * f2[:, t2] * f3[:, t1])
    return s / 6.0


def get_cond(i, msk, QQ_tilde):
    """cmbbest.pyx:50-53."""
    m = msk.copy()
    m[i] = True
    return np.linalg.cond(QQ_tilde[np.ix_(m, m)])


def basis_expansion(grid, S, regularization=0, Niter=1e4, rtol=1e-12, bad_modes=False, RR_tol=1e-2,
                    cond_tol=1e9, max_modes=None, QQ=None):
    """Returns (alpha, shape_covariance, convergence_correlation, convergence_MSE, masks)."""
    S = np.atleast_2d(np.asarray(S, dtype=np.float64))
    p1, p2, p3 = (np.asarray(grid[k]) for k in ("p1", "p2", "p3"))
    w = np.asarray(grid["w"])
    N_models, N_modes = S.shape[0], len(p1)
    Q = mode_bispectra(grid)
    if QQ is None:
        QQ = pyoracle.mode_covariance(grid)
    norms = np.sqrt(np.diag(QQ))
    if max_modes is None:
        max_modes = p1.size
    shape_covariance = np.matmul(S * w[np.newaxis, :], S.T)
    QS = np.matmul(S * w[np.newaxis, :], Q.T)
    QQ_temp, QS_temp, Q_temp = QQ * 1, QS * 1, Q * 1
    badmask = (((p1 <= 2) & (p2 == 0) & (p3 == 0)) | ((p1 == 1) & (p2 == 1) & (p3 == 0)))
    badmask |= bad_modes
    badinds = np.where(badmask)[0]
    for bi in badinds:
        QQ_temp[bi, :] = 0
        QQ_temp[:, bi] = 0
        QQ_temp[bi, bi] = QQ[bi, bi]
        QS_temp[:, bi] = 0
        Q_temp[bi] = 0
    QQ_tilde = QQ_temp / norms[:, np.newaxis] / norms[np.newaxis, :]
    Q_tilde = Q_temp / norms[:, np.newaxis]
    alpha = np.zeros((N_models, N_modes))
    masks = np.zeros((N_models, N_modes), bool)
    for model_no in range(N_models):
        QS_tilde = QS_temp[model_no, :] / norms
        OMP_msk = np.zeros_like(p1, dtype=bool)
        OMP_msk[(p1 <= 3) & (p2 <= 3) & (p3 <= 3)] = True
        if max_modes == -1:
            OMP_msk = np.ones_like(OMP_msk, dtype=bool)
        while True:
            if len(norms) <= np.sum(OMP_msk):
                break
            A = QQ_tilde[np.ix_(OMP_msk, OMP_msk)] + np.eye(np.sum(OMP_msk)) * regularization
            alpha_tmp, _ = conjugate_gradient(A, QS_tilde[OMP_msk], rtol=rtol, atol=0, maxiter=int(Niter))
            R = S - np.sum(Q_tilde[OMP_msk] * alpha_tmp[:, None], axis=0)
            RR = np.matmul(R * w[np.newaxis, :], R.T)[0, 0]
            delta_RR = RR / shape_covariance
            if np.abs(delta_RR[0, 0]) <= RR_tol:
                break
            if np.sum(OMP_msk) == max_modes - 1:
                break
            cond_here = get_cond(np.where(OMP_msk)[0][0], OMP_msk, QQ_tilde)
            cond = np.array([get_cond(i, OMP_msk, QQ_tilde) for i in np.where(np.invert(OMP_msk))[0]])
            fcond = cond / cond_here - 1
            if cond_here > cond_tol:
                break
            QR = np.matmul(R * w[np.newaxis, :], Q.T)[0] / norms / np.sqrt(RR)  # compute_QS_C(R)[0]
            QR[badinds] = 0
            QR = QR[np.invert(OMP_msk)]
            m = np.argsort(np.abs(QR))[-10:]
            m = m[np.argmin(fcond[m])]
            m = np.where(np.invert(OMP_msk))[0][m]
            OMP_msk[m] = True
        A = QQ_tilde[np.ix_(OMP_msk, OMP_msk)] + np.eye(np.sum(OMP_msk)) * regularization
        alpha_tmp, _ = conjugate_gradient(A, QS_tilde[OMP_msk], rtol=1e-16, atol=0, maxiter=int(1e6))
        alpha_tilde = np.zeros_like(norms)
        alpha_tilde[OMP_msk] = alpha_tmp
        alpha[model_no, :] = alpha_tilde / norms
        masks[model_no] = OMP_msk
    sum_SS = np.diag(shape_covariance)
    sum_SR = np.sum(alpha * QS, axis=1)
    sum_RR = np.sum(alpha * np.matmul(alpha, QQ), axis=1)
    corr = sum_SR / np.sqrt(sum_SS * sum_RR)
    corr = 1.0 - np.abs(1 - corr)
    mse = np.abs(sum_SS + sum_RR - 2 * sum_SR) / sum_SS
    return alpha, shape_covariance, corr, mse, masks
