"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Arrays use the
reference's layouts: complex spectra as (N, N, N/2+1) complex128, padded real
fields as (N, N, 2*(N/2+1)) float64 (Aarambam/LPT/Basis/main.c:456, 772).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_INT_FIELDS = ["Nmesh", "Nsample", "Nmax", "N_modes", "Seed", "SphereMode", "FixedAmplitude",
               "PhaseFlip", "WhichSpectrum", "WhichTransfer", "GlassTileFac", "WDM_On"]
_DBL_FIELDS = ["Box", "Omega", "OmegaLambda", "OmegaBaryon", "HubbleParam", "Sigma8",
               "PrimordialIndex", "Redshift", "Fnl", "ShapeGamma", "UnitLength_in_cm",
               "UnitMass_in_g", "UnitVelocity_in_cm_per_s", "InputSpectrum_UnitLength_in_cm",
               "WDM_PartMass_in_kev"]


class OrcParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS] + [(n, C.c_double) for n in _DBL_FIELDS] +
                [("ntransfer", C.c_int), ("transfer_k", C.POINTER(C.c_double)),
                 ("transfer_t", C.POINTER(C.c_double)), ("Variant", C.c_int)] +
                [(n, C.c_double) for n in ["A_res", "w_res", "phi_res"]])


class OrcCosmo(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ["UnitTime_in_s", "G", "Hubble", "InitTime", "Norm", "sigma8_in", "D0", "Dplus",
                 "Anorm", "Beta", "hubble_a", "vel_prefac", "vel_prefac2", "fac"]]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        P = C.POINTER(OrcParams)
        L.orc_ranlxd_nth_get.restype = C.c_ulong
        L.orc_ranlxd_nth_get.argtypes = [C.c_ulong, C.c_int, C.c_int]
        L.orc_ranlxd_draws.argtypes = [C.c_ulong, C.c_int, C.c_int, dp]
        L.orc_legendre_Pl.restype = C.c_double
        L.orc_legendre_Pl.argtypes = [C.c_int, C.c_double]
        L.orc_seedtable.argtypes = [C.c_int, C.c_int, up]
        L.orc_cosmology.argtypes = [P, C.POINTER(OrcCosmo)]
        L.orc_transfer.restype = C.c_double
        L.orc_transfer.argtypes = [P, C.c_double]
        L.orc_qromb_growth.restype = C.c_double
        L.orc_qromb_growth.argtypes = [P, C.c_double]
        L.orc_sigma2_tophat.restype = C.c_double
        L.orc_sigma2_tophat.argtypes = [P, C.c_double, C.c_double]
        L.orc_alpha_symmetrise.argtypes = [C.c_int, C.c_int, ip, ip, ip, dp, dp]
        L.orc_c2r.argtypes = [C.c_int, dp]
        L.orc_r2c.argtypes = [C.c_int, dp]
        L.orc_gaussian.argtypes = [P, C.POINTER(OrcCosmo), up, dp]
        L.orc_ng_potential.argtypes = [P, C.POINTER(OrcCosmo), dp, dp, dp, C.c_void_p, C.c_void_p]
        L.orc_lpt_fields.argtypes = [P, C.POINTER(OrcCosmo), dp] + [dp] * 6
        L.orc_glass_tile.argtypes = [P, fp, C.c_int, C.c_double, fp, C.c_void_p]
        L.orc_particles.argtypes = [P, C.POINTER(OrcCosmo), C.c_longlong, C.c_void_p, C.c_void_p,
                                    fp, fp]
        L.orc_wdm_thermal.argtypes = [P, C.c_longlong, fp]
        L.orc_drand48.restype = C.c_double
        L.orc_drand48.argtypes = []
        L.orc_drand48_reset.argtypes = []
        L.orc_run.argtypes = [P, dp, dp, fp, C.c_int, C.c_double, fp, fp, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        modal = [dp, ip, ip, ip, C.c_int, ip, ip, ip, C.c_int, dp, C.c_int, C.c_int]
        L.orc_mode_covariance.argtypes = [dp] + modal
        L.orc_mode_qs.argtypes = [dp, dp, C.c_int] + modal
        _lib = L
    return _lib


REF_PATH = os.path.join(HERE, "_ref", "libtetrapyd_ref.so")
_ref = None


def ref_lib():
    """The reference's own tetrapyd.c/arrays.c compiled by `make -C oracle ref` (None if not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        modal = [dp, ip, ip, ip, C.c_int, ip, ip, ip, C.c_int, dp, C.c_int, C.c_int]
        L.compute_mode_bispectra_covariance.argtypes = [dp] + modal
        L.compute_QS.argtypes = [dp, dp, C.c_int] + modal
        _ref = L
    return _ref


def modal_args(grid):
    """Argument tuple (w, i1, i2, i3, npts, p1, p2, p3, n_modes, evals, pmax, nk) of a modal grid dict."""
    c = lambda a, t: np.ascontiguousarray(a, dtype=t)
    return (c(grid["w"], np.float64), c(grid["i1"], np.int32), c(grid["i2"], np.int32), c(grid["i3"], np.int32),
            len(grid["w"]), c(grid["p1"], np.int32), c(grid["p2"], np.int32), c(grid["p3"], np.int32),
            len(grid["p1"]), c(grid["evals"], np.float64).reshape(-1), grid["evals"].shape[0], grid["evals"].shape[1])


def mode_covariance(grid, use_ref=False):
    a = modal_args(grid)
    out = np.zeros(a[8] * a[8])
    L = ref_lib() if use_ref else lib()
    (L.compute_mode_bispectra_covariance if use_ref else L.orc_mode_covariance)(out, *a)
    return out.reshape(a[8], a[8])


def mode_qs(grid, S, use_ref=False):
    a = modal_args(grid)
    S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, a[4])
    out = np.zeros(S.shape[0] * a[8])
    L = ref_lib() if use_ref else lib()
    (L.compute_QS if use_ref else L.orc_mode_qs)(out, S.reshape(-1), S.shape[0], *a)
    return out.reshape(S.shape[0], a[8])


def make_params(cfg):
    """cfg: dict with the reference's parameter-file tags (+ transfer_k/transfer_t arrays)."""
    p = OrcParams()
    for n in _INT_FIELDS:
        setattr(p, n, int(cfg.get(n, 0)))
    for n in _DBL_FIELDS + ["A_res", "w_res", "phi_res"]:
        setattr(p, n, float(cfg.get(n, 0.0)))
    p.Variant = int(cfg.get("Variant", 0))
    keep = []
    if cfg.get("WhichTransfer", 0) == 2:
        tk = np.ascontiguousarray(cfg["transfer_k"], dtype=np.float64)
        tt = np.ascontiguousarray(cfg["transfer_t"], dtype=np.float64)
        keep = [tk, tt]
        p.ntransfer = len(tk)
        p.transfer_k = tk.ctypes.data_as(C.POINTER(C.c_double))
        p.transfer_t = tt.ctypes.data_as(C.POINTER(C.c_double))
    p._keep = keep
    return p


def ranlxd_nth_get(seed, luxury, n):
    return lib().orc_ranlxd_nth_get(seed, luxury, n)


def ranlxd_draws(seed, n, luxury=202):
    out = np.empty(n, np.float64)
    lib().orc_ranlxd_draws(seed, luxury, n, out)
    return out


def legendre(l, x):
    return lib().orc_legendre_Pl(l, x)


def seedtable(N, seed):
    out = np.empty(N * N, np.uint32)
    lib().orc_seedtable(N, seed, out)
    return out.reshape(N, N)


def cosmology(cfg):
    p = make_params(cfg)
    c = OrcCosmo()
    rc = lib().orc_cosmology(C.byref(p), C.byref(c))
    if rc:
        raise RuntimeError(f"oracle cosmology failed with code {rc}")
    return {n: getattr(c, n) for n, _ in OrcCosmo._fields_}


def transfer(cfg, k):
    p = make_params(cfg)
    return lib().orc_transfer(C.byref(p), k)


def alpha_table(Nm, rows):
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 4)
    n1, n2, n3 = (np.ascontiguousarray(rows[:, c].astype(np.int32)) for c in range(3))
    a = np.ascontiguousarray(rows[:, 3])
    out = np.zeros(Nm ** 3, np.float64)
    lib().orc_alpha_symmetrise(Nm, len(a), n1, n2, n3, a, out)
    return out.reshape(Nm, Nm, Nm)


def c2r(spec):
    """spec (N,N,N/2+1) complex -> real (N,N,N) with FFTW2 rfftwnd semantics (unnormalised)."""
    N = spec.shape[0]
    buf = np.ascontiguousarray(spec.astype(np.complex128)).view(np.float64).reshape(-1).copy()
    lib().orc_c2r(N, buf)
    return buf.reshape(N, N, 2 * (N // 2 + 1))[:, :, :N].copy()


def r2c(real):
    N = real.shape[0]
    pad = np.zeros((N, N, 2 * (N // 2 + 1)), np.float64)
    pad[:, :, :N] = real
    buf = pad.reshape(-1)
    lib().orc_r2c(N, buf)
    return buf.view(np.complex128).reshape(N, N, N // 2 + 1).copy()


def gaussian(cfg, st=None):
    N = cfg["Nmesh"]
    p = make_params(cfg)
    c = OrcCosmo()
    rc = lib().orc_cosmology(C.byref(p), C.byref(c))
    if rc:
        raise RuntimeError(rc)
    if st is None:
        st = seedtable(N, cfg["Seed"])
    out = np.empty(2 * N * N * (N // 2 + 1), np.float64)
    lib().orc_gaussian(C.byref(p), C.byref(c), np.ascontiguousarray(st.reshape(-1)), out)
    return out.view(np.complex128).reshape(N, N, N // 2 + 1)


def run(cfg, alpha, x0, glass, glass_box, want_fields=True):
    """Full displacement_fields() for one rank.  Returns dict of outputs."""
    N = cfg["Nmesh"]
    T = cfg["GlassTileFac"]
    p = make_params(cfg)
    glass = np.ascontiguousarray(glass, dtype=np.float32).reshape(-1)
    ng = glass.size // 3
    npart = T ** 3 * ng
    pos = np.empty(3 * npart, np.float32)
    vel = np.empty(3 * npart, np.float32)
    ids = np.empty(npart, np.int32)
    nd = 2 * N * N * (N // 2 + 1)
    outs = {}
    ptrs = []
    for name in ["phig_spec", "phi_spec", "phig_real", "phing_real"]:
        if want_fields:
            outs[name] = np.empty(nd, np.float64)
            ptrs.append(outs[name].ctypes.data_as(C.c_void_p))
        else:
            ptrs.append(None)
    rc = lib().orc_run(C.byref(p), np.ascontiguousarray(alpha.reshape(-1), dtype=np.float64),
                       np.ascontiguousarray(x0.reshape(-1), dtype=np.float64), glass, ng,
                       float(glass_box), pos, vel, ids.ctypes.data_as(C.c_void_p), *ptrs)
    if rc:
        raise RuntimeError(f"oracle run failed with code {rc}")
    res = {"pos": pos.reshape(-1, 3), "vel": vel.reshape(-1, 3), "ids": ids}
    H = N // 2 + 1
    if want_fields:
        res["phig_spec"] = outs["phig_spec"].view(np.complex128).reshape(N, N, H)
        res["phi_spec"] = outs["phi_spec"].view(np.complex128).reshape(N, N, H)
        res["phig_real"] = outs["phig_real"].reshape(N, N, 2 * H)[:, :, :N]
        res["phing_real"] = outs["phing_real"].reshape(N, N, 2 * H)[:, :, :N]
    return res


def lpt_fields(cfg, phi_spec):
    """ZA and 2LPT real displacement fields (main.c:1158-1381) from Phi(k).

    Returns six (N,N,N) arrays: disp x,y,z and disp2 x,y,z (disp2 not divided by N^3)."""
    N = cfg["Nmesh"]
    p = make_params(cfg)
    c = OrcCosmo()
    rc = lib().orc_cosmology(C.byref(p), C.byref(c))
    if rc:
        raise RuntimeError(rc)
    nd = 2 * N * N * (N // 2 + 1)
    buf = np.ascontiguousarray(phi_spec, dtype=np.complex128).view(np.float64).reshape(-1)
    outs = [np.empty(nd, np.float64) for _ in range(6)]
    lib().orc_lpt_fields(C.byref(p), C.byref(c), buf, *outs)
    return [o.reshape(N, N, 2 * (N // 2 + 1))[:, :, :N].copy() for o in outs]


def glass_tile(cfg, glass, glass_box, with_ids=False):
    p = make_params(cfg)
    glass = np.ascontiguousarray(glass, dtype=np.float32).reshape(-1)
    ng = glass.size // 3
    npart = cfg["GlassTileFac"] ** 3 * ng
    pos = np.empty(3 * npart, np.float32)
    ids = np.zeros(npart, np.int32)
    lib().orc_glass_tile(C.byref(p), glass, ng, float(glass_box), pos, ids.ctypes.data_as(C.c_void_p))
    return (pos.reshape(-1, 3), ids) if with_ids else pos.reshape(-1, 3)


def particles(cfg, pos, disp, disp2):
    """CIC readout and update (main.c:1440-1516) of float32 positions (n, 3) from six real
    (N, N, N) fields: disp x, y, z and disp2 x, y, z (disp2 not divided by N^3).
    Returns (new positions, velocities)."""
    N = cfg["Nmesh"]
    p = make_params(cfg)
    c = OrcCosmo(**cosmology(cfg))
    pad = []
    for f in list(disp) + list(disp2):
        a = np.zeros((N, N, 2 * (N // 2 + 1)), np.float64)
        a[:, :, :N] = f
        pad.append(a)
    ptrs = lambda fs: (C.c_void_p * 3)(*[f.ctypes.data for f in fs])
    P = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1).copy()
    V = np.zeros_like(P)
    lib().orc_particles(C.byref(p), C.byref(c), P.size // 3, ptrs(pad[:3]), ptrs(pad[3:]), P, V)
    return P.reshape(-1, 3), V.reshape(-1, 3)


def wdm_thermal(cfg, vel):
    """Velocities as the snapshot writer writes them with WDM_On 1 and WDM_Vtherm_On 1
    (save.c:350-356, power.c:624-707), drand48 from its default state."""
    p = make_params(cfg)
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1).copy()
    lib().orc_wdm_thermal(C.byref(p), v.size // 3, v)
    return v.reshape(-1, 3)


def drand48(n):
    """n values of the oracle's drand48 from its default state."""
    L = lib()
    L.orc_drand48_reset()
    return np.array([L.orc_drand48() for _ in range(n)])
