"""ctypes binding of libaarambam_hip.so (include/aarambam.h).

Fails loudly when the library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaarambam_hip.so")

AAR_FP64, AAR_FP32 = 64, 32
AAR_SCHED_PLAIN, AAR_SCHED_BATCHED = 0, 1
FIELDS = dict(phig_spec=0, phi_spec=1, disp_x=2, disp_y=3, disp_z=4, disp2_x=5, disp2_y=6, disp2_z=7,
              phig_real=8, phing_real=9)

_INTS = ["Nmesh", "Nsample", "Nmax", "N_modes", "Seed", "SphereMode", "FixedAmplitude", "PhaseFlip",
         "SavePotentialField", "WhichSpectrum", "WhichTransfer", "GlassTileFac",
         "NumFilesWrittenInParallel", "WDM_On", "WDM_Vtherm_On"]
_DBLS = ["Box", "Omega", "OmegaLambda", "OmegaBaryon", "OmegaDM_2ndSpecies", "HubbleParam", "Sigma8",
         "PrimordialIndex", "Redshift", "Fnl", "ShapeGamma", "UnitLength_in_cm", "UnitMass_in_g",
         "UnitVelocity_in_cm_per_s", "InputSpectrum_UnitLength_in_cm", "WDM_PartMass_in_kev"]
_STRS = ["GlassFile", "FileWithInputSpectrum", "FileWithInputTransfer", "FileWithAlphaCoeff",
         "FileWithX0Coeff", "OutputDir", "FileBase"]


_RES_INT = ["Variant"]  # 0 Basis, 1 ResBasis
_RES_DBL = ["A_res", "w_res", "phi_res"]


class AarParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _INTS] + [(n, C.c_double) for n in _DBLS] +
                [(n, C.c_char * 500) for n in _STRS] + [(n, C.c_int) for n in _RES_INT] +
                [(n, C.c_double) for n in _RES_DBL])


class AarTables(C.Structure):
    _fields_ = [("alpha", C.POINTER(C.c_double)), ("x0", C.POINTER(C.c_double)),
                ("ntransfer", C.c_int), ("transfer_k", C.POINTER(C.c_double)),
                ("transfer_t", C.POINTER(C.c_double)), ("glass", C.POINTER(C.c_float)),
                ("nglass", C.c_int), ("glass_box", C.c_double), ("glass_header", C.POINTER(C.c_uint8))]


class AarCosmo(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ["UnitTime_in_s", "G", "Hubble", "InitTime", "Norm", "sigma8_in", "D0", "Dplus", "Anorm",
                 "Beta", "hubble_a", "vel_prefac", "vel_prefac2", "fac"]]


EXPORTS = ["aar_status_string", "aar_read_params", "aar_read_params_res", "aar_load_alpha", "aar_load_x0", "aar_load_transfer",
           "aar_read_glass", "aar_cosmology", "aar_seedtable", "aar_seedtable_threads", "aar_ranlxd_nth", "aar_write_gadget",
           "aar_write_gadget_files", "aar_task_layout", "aar_write_gadget_tasks", "aar_write_phi_tasks",
           "aar_write_phi", "aar_write_inputspec", "aar_ctx_create", "aar_ctx_destroy", "aar_ctx_sync", "aar_job_bytes",
           "aar_job_prepare", "aar_job_run", "aar_job_sync", "aar_job_device_output", "aar_job_fetch",
           "aar_job_fetch_field", "aar_job_last_seconds", "aar_job_destroy", "aar_fft3d",
           "aar_gaussian_field", "aar_ranlxd_device", "aar_bench_pass", "aar_job_set_profiling", "aar_job_kernel_stats", "aar_job_set_graph",
           "aar_server_create", "aar_server_submit", "aar_server_wait", "aar_server_fetch",
           "aar_server_device_output", "aar_server_release", "aar_server_stats", "aar_server_destroy",
           "aar_autotune", "aar_xy_tile_variant", "aar_set_xy_tile_variant", "aar_xy_group_cols",
           "aar_set_xy_group_cols", "aar_job_fetch_field_box",
           "aar_job_ng_statistics", "aar_mode_bispectra_covariance", "aar_mode_QS",
           "aar_mode_basis_expansion", "aar_job_reseed", "aar_ctx_cached_plans",
           "aar_job_run_noalloc", "aar_autotune_plans", "aar_xy_radix_plan", "aar_set_xy_radix_plan",
           "aar_sym_eig_extremes", "aar_ctx_set_jit", "aar_job_jit", "aar_autotune_report",
           "aar_xy_split_group_cols", "aar_set_xy_split_group_cols", "aar_ctx_set_modes_x"]
KINDS = ["gauss", "xy_pass", "z_pass", "modes_spec", "products", "combine", "lpt_spec", "cic", "accumulate",
         "xy_split", "modes_xy"]


class AarError(RuntimeError):
    def __init__(self, code, what=""):
        msg = lib().aar_status_string(code).decode() if _lib is not None else ""
        super().__init__(f"{what}: status {code} ({msg})")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m aarambam_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, dp, fp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        PP, TP, CP = C.POINTER(AarParams), C.POINTER(AarTables), C.POINTER(AarCosmo)
        sig = {
            "aar_status_string": (C.c_char_p, [C.c_int]),
            "aar_read_params": (C.c_int, [C.c_char_p, PP, C.c_char_p, C.c_size_t]),
            "aar_read_params_res": (C.c_int, [C.c_char_p, PP, C.c_char_p, C.c_size_t]),
            "aar_load_alpha": (C.c_int, [C.c_char_p, C.c_int, dp]),
            "aar_load_x0": (C.c_int, [C.c_char_p, C.c_int, dp]),
            "aar_load_transfer": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), dp, dp]),
            "aar_read_glass": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), fp, dp, C.POINTER(C.c_uint8)]),
            "aar_cosmology": (C.c_int, [PP, TP, CP]),
            "aar_seedtable": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
            "aar_seedtable_threads": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_uint32), C.c_int]),
            "aar_ranlxd_nth": (C.c_ulong, [C.c_ulong, C.c_int, C.c_int]),
            "aar_write_gadget": (C.c_int, [PP, CP, fp, fp, C.c_longlong, C.POINTER(C.c_uint8)]),
            "aar_write_gadget_files": (C.c_int, [PP, CP, fp, fp, C.c_longlong, C.POINTER(C.c_uint8), C.c_int,
                                                 C.c_longlong, C.POINTER(C.c_int)]),
            "aar_write_phi": (C.c_int, [PP, dp, C.c_int]),
            "aar_task_layout": (C.c_int, [PP, TP, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
            "aar_write_gadget_tasks": (C.c_int, [PP, CP, TP, fp, fp, C.c_longlong, C.c_int]),
            "aar_write_phi_tasks": (C.c_int, [PP, dp, C.c_int, C.c_int, C.c_int]),
            "aar_write_inputspec": (C.c_int, [PP, TP, CP]),
            "aar_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "aar_ctx_destroy": (C.c_int, [vp]),
            "aar_ctx_sync": (C.c_int, [vp]),
            "aar_job_bytes": (C.c_longlong, [PP, TP, C.c_int, C.c_int]),
            "aar_job_prepare": (C.c_int, [vp, PP, TP, C.c_int, C.c_int, C.POINTER(vp)]),
            "aar_job_run": (C.c_int, [vp, vp]),
            "aar_job_reseed": (C.c_int, [vp, C.c_int, C.c_double]),
            "aar_job_run_noalloc": (C.c_int, [vp, PP, TP, C.c_int, C.c_int, vp, vp, vp]),
            "aar_ctx_cached_plans": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
            "aar_job_sync": (C.c_int, [vp]),
            "aar_job_device_output": (C.c_int, [vp, C.POINTER(fp), C.POINTER(fp), C.POINTER(C.c_longlong)]),
            "aar_job_fetch": (C.c_int, [vp, fp, fp]),
            "aar_job_fetch_field": (C.c_int, [vp, C.c_int, dp]),
            "aar_job_fetch_field_box": (C.c_int, [vp, C.c_int] + [C.c_int] * 6 + [dp]),
            "aar_job_ng_statistics": (C.c_int, [vp, C.c_int] + [dp] * 8),
            "aar_mode_bispectra_covariance": (C.c_int, [vp, dp, dp, ip, ip, ip, C.c_int, ip, ip, ip, C.c_int, dp,
                                                        C.c_int, C.c_int]),
            "aar_mode_QS": (C.c_int, [vp, dp, dp, C.c_int, dp, ip, ip, ip, C.c_int, ip, ip, ip, C.c_int, dp, C.c_int,
                                      C.c_int]),
            "aar_mode_basis_expansion": (C.c_int, [vp, dp, dp, dp, dp, C.POINTER(C.c_uint8), dp, C.c_int, dp, ip, ip,
                                                   ip, C.c_int, ip, ip, ip, C.c_int, dp, C.c_int, C.c_int,
                                                   C.POINTER(C.c_uint8), C.c_double, C.c_longlong, C.c_double,
                                                   C.c_double, C.c_double, C.c_int]),
            "aar_job_last_seconds": (C.c_double, [vp]),
            "aar_job_destroy": (C.c_int, [vp]),
            "aar_fft3d": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp]),
            "aar_gaussian_field": (C.c_int, [vp, PP, TP, C.c_int, dp]),
            "aar_ranlxd_device": (C.c_int, [vp, C.c_uint32, C.c_int, dp]),
            "aar_sym_eig_extremes": (C.c_int, [vp, C.c_int, C.c_int, dp, dp]),
            "aar_bench_pass": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
            "aar_job_set_profiling": (C.c_int, [vp, C.c_int]),
            "aar_job_set_graph": (C.c_int, [vp, C.c_int]),
            "aar_ctx_set_jit": (C.c_int, [vp, C.c_int]),
            "aar_ctx_set_modes_x": (C.c_int, [vp, C.c_int]),
            "aar_job_jit": (C.c_int, [vp]),
            "aar_job_kernel_stats": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), dp, dp]),
            "aar_server_create": (C.c_int, [vp, C.c_longlong, C.c_int, C.POINTER(vp)]),
            "aar_server_submit": (C.c_int, [vp, PP, TP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong)]),
            "aar_server_wait": (C.c_int, [vp, C.c_longlong, dp, dp]),
            "aar_server_fetch": (C.c_int, [vp, C.c_longlong, fp, fp]),
            "aar_server_device_output": (C.c_int, [vp, C.c_longlong, C.POINTER(fp), C.POINTER(fp),
                                                   C.POINTER(C.c_longlong)]),
            "aar_server_release": (C.c_int, [vp, C.c_longlong]),
            "aar_server_stats": (C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                           C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
            "aar_server_destroy": (C.c_int, [vp]),
            "aar_autotune": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, C.POINTER(C.c_int)]),
            "aar_xy_tile_variant": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "aar_autotune_plans": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, C.POINTER(C.c_int)]),
            "aar_xy_radix_plan": (C.c_int, [C.c_int, C.c_int]),
            "aar_set_xy_radix_plan": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "aar_set_xy_tile_variant": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "aar_xy_group_cols": (C.c_int, [C.c_int, C.c_int]),
            "aar_set_xy_group_cols": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "aar_autotune_report": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp]),
            "aar_xy_split_group_cols": (C.c_int, [C.c_int, C.c_int]),
            "aar_set_xy_split_group_cols": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise AarError(rc, what)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params_from_dict(cfg):
    p = AarParams()
    for n in _INTS + _RES_INT:
        setattr(p, n, int(cfg.get(n, 0)))
    for n in _DBLS + _RES_DBL:
        setattr(p, n, float(cfg.get(n, 0.0)))
    for n in _STRS:
        setattr(p, n, str(cfg.get(n, "")).encode())
    return p


def params_to_dict(p):
    d = {n: getattr(p, n) for n in _INTS + _DBLS + _RES_INT + _RES_DBL}
    d.update({n: getattr(p, n).decode() for n in _STRS})
    return d


def read_params(path, res=False):
    """Parse a parameter file: the Basis tag set, or with res=True the ResBasis one."""
    p = AarParams()
    msg = C.create_string_buffer(4096)
    fn = lib().aar_read_params_res if res else lib().aar_read_params
    rc = fn(path.encode(), C.byref(p), msg, 4096)
    if rc:
        raise AarError(rc, msg.value.decode())
    return p


class tables:
    """Holds the host arrays of one job and the aar_tables view of them."""

    def __init__(self, alpha, x0, glass, glass_box, transfer_k=None, transfer_t=None, glass_header=None):
        self.alpha = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1)
        self.x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
        self.glass = np.ascontiguousarray(glass, dtype=np.float32).reshape(-1)
        self.tk = None if transfer_k is None else np.ascontiguousarray(transfer_k, dtype=np.float64)
        self.tt = None if transfer_t is None else np.ascontiguousarray(transfer_t, dtype=np.float64)
        self.hdr = None if glass_header is None else np.ascontiguousarray(glass_header, dtype=np.uint8)
        t = AarTables()
        t.alpha = _dptr(self.alpha)
        t.x0 = _dptr(self.x0)
        t.glass = self.glass.ctypes.data_as(C.POINTER(C.c_float))
        t.nglass = self.glass.size // 3
        t.glass_box = float(glass_box)
        if self.tk is not None:
            t.ntransfer = len(self.tk)
            t.transfer_k = _dptr(self.tk)
            t.transfer_t = _dptr(self.tt)
        if self.hdr is not None:
            t.glass_header = self.hdr.ctypes.data_as(C.POINTER(C.c_uint8))
        self.c = t

    @property
    def nglass(self):
        return self.glass.size // 3


def cosmology(params, tabs):
    c = AarCosmo()
    _check(lib().aar_cosmology(C.byref(params), C.byref(tabs.c), C.byref(c)), "aar_cosmology")
    return {n: getattr(c, n) for n, _ in AarCosmo._fields_}


def seedtable(N, seed, threads=None):
    """The reference's N x N per-pencil seed table; threads: host threads that fill it (default:
    the library's choice), bit-identical for every count."""
    out = np.empty(N * N, np.uint32)
    ptr = out.ctypes.data_as(C.POINTER(C.c_uint32))
    if threads is None:
        _check(lib().aar_seedtable(N, seed, ptr), "aar_seedtable")
    else:
        _check(lib().aar_seedtable_threads(N, seed, ptr, threads), "aar_seedtable_threads")
    return out.reshape(N, N)


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().aar_ctx_create(device, C.byref(h)), "aar_ctx_create")
        self.h = h

    def sync(self):
        _check(lib().aar_ctx_sync(self.h), "aar_ctx_sync")

    def set_jit(self, on=True):
        """Batched jobs prepared from now on compile their fused z pass for their own product
        schedule (on) or run the library's kernel (off); results are identical."""
        _check(lib().aar_ctx_set_jit(self.h, int(on)), "aar_ctx_set_jit")

    def set_modes_x(self, mode=1):
        """Batched jobs run from now on filter their mode fields in the loads of their x pass when
        the filters are short (1, the default), always (2) or never (0, the mode fields written
        in k-space first); results are identical."""
        _check(lib().aar_ctx_set_modes_x(self.h, int(mode)), "aar_ctx_set_modes_x")

    def close(self):
        if self.h:
            lib().aar_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fft3d(self, data, dir, precision=AAR_FP64):
        data = np.asarray(data)
        N = data.shape[0]
        if dir > 0:
            inp = np.ascontiguousarray(data, dtype=np.complex128).view(np.float64).reshape(-1)
            out = np.empty(N ** 3, np.float64)
        else:
            inp = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
            out = np.empty(2 * N * N * (N // 2 + 1), np.float64)
        _check(lib().aar_fft3d(self.h, N, precision, dir, _dptr(inp), _dptr(out)), "aar_fft3d")
        if dir > 0:
            return out.reshape(N, N, N)
        return out.view(np.complex128).reshape(N, N, N // 2 + 1)

    def gaussian_field(self, params, tabs, precision=AAR_FP64):
        N = params.Nmesh
        out = np.empty(2 * N * N * (N // 2 + 1), np.float64)
        _check(lib().aar_gaussian_field(self.h, C.byref(params), C.byref(tabs.c), precision, _dptr(out)),
               "aar_gaussian_field")
        return out.view(np.complex128).reshape(N, N, N // 2 + 1)

    def bench_pass(self, n, kind, nf=4, iters=5, precision=AAR_FP32):
        s = C.c_double(0)
        _check(lib().aar_bench_pass(self.h, n, precision, kind, nf, iters, C.byref(s)), "aar_bench_pass")
        return s.value

    def autotune_plans(self, n, precision=AAR_FP32, nf=8):
        """aar_autotune with the radix plans' timings: (seconds per plan [3], chosen plan)."""
        sec = np.full(3, -1.0)
        pl = C.c_int(0)
        _check(lib().aar_autotune_plans(self.h, n, precision, nf, _dptr(sec), C.byref(pl)), "aar_autotune_plans")
        return sec, pl.value

    def autotune_report(self, n, precision=AAR_FP32, nf=8):
        """One autotuner search with every timing (seconds, -1 where not timed): tile shapes [2][4],
        x + y pass pairs per radix plan and column group [3][4] (whole passes, all columns, halves,
        quarters), multi-output pairs [9]: per column group (the same and eighths), then per tile
        shape of their y pass at the group kept."""
        ts, ps, ss = np.full(8, -1.0), np.full(12, -1.0), np.full(9, -1.0)
        _check(lib().aar_autotune_report(self.h, n, precision, nf, _dptr(ts), _dptr(ps), _dptr(ss)),
               "aar_autotune_report")
        return ts.reshape(2, 4), ps.reshape(3, 4), ss

    def cached_plans(self):
        """(count, device bytes) of the job tables the context keeps resident for later jobs."""
        n, b = C.c_int(0), C.c_longlong(0)
        _check(lib().aar_ctx_cached_plans(self.h, C.byref(n), C.byref(b)), "aar_ctx_cached_plans")
        return n.value, b.value

    def autotune(self, n, precision=AAR_FP32, nf=8):
        """Time the x / y tile shapes at n^3 and keep the fastest; returns (seconds[2][4], chosen[2])."""
        sec = np.zeros(8, np.float64)
        ch = (C.c_int * 2)()
        _check(lib().aar_autotune(self.h, n, precision, nf, _dptr(sec), ch), "aar_autotune")
        return sec.reshape(2, 4), [ch[0], ch[1]]

    def _modal(self, grid):
        c = lambda a, t: np.ascontiguousarray(a, dtype=t)
        ev = c(grid["evals"], np.float64)
        arrs = dict(w=c(grid["w"], np.float64), i1=c(grid["i1"], np.int32), i2=c(grid["i2"], np.int32),
                    i3=c(grid["i3"], np.int32), p1=c(grid["p1"], np.int32), p2=c(grid["p2"], np.int32),
                    p3=c(grid["p3"], np.int32), ev=ev)
        ipt = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        args = [_dptr(arrs["w"]), ipt(arrs["i1"]), ipt(arrs["i2"]), ipt(arrs["i3"]), len(arrs["w"]),
                ipt(arrs["p1"]), ipt(arrs["p2"]), ipt(arrs["p3"]), len(arrs["p1"]), _dptr(ev), ev.shape[0],
                ev.shape[1]]
        return arrs, args

    def mode_covariance(self, grid):
        """Mode-bispectrum covariance on the device (tetrapyd.c:111-208); grid: dict of
        w, i1, i2, i3 (tetrapyd points), p1, p2, p3 (3-D modes), evals (pmax x nk)."""
        keep, args = self._modal(grid)
        n = args[8]
        out = np.zeros(n * n)
        _check(lib().aar_mode_bispectra_covariance(self.h, _dptr(out), *args), "aar_mode_bispectra_covariance")
        return out.reshape(n, n)

    def mode_qs(self, grid, S):
        """Projections of shapes S (n_shapes x npts) on the modes (tetrapyd.c:210-281)."""
        keep, args = self._modal(grid)
        S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, args[4])
        out = np.zeros(S.shape[0] * args[8])
        _check(lib().aar_mode_QS(self.h, _dptr(out), _dptr(S), S.shape[0], *args), "aar_mode_QS")
        return out.reshape(S.shape[0], args[8])

    def sym_eig_extremes(self, mats):
        """(min |lambda|, max |lambda|) of each symmetric matrix of mats (nmat x m x m), from the
        condition-number kernel of the basis expansion."""
        mats = np.ascontiguousarray(mats, dtype=np.float64)
        if mats.ndim == 2:
            mats = mats[None]
        nmat, m = mats.shape[0], mats.shape[1]
        ext = np.zeros((nmat, 2))
        _check(lib().aar_sym_eig_extremes(self.h, m, nmat, _dptr(mats), _dptr(ext)), "aar_sym_eig_extremes")
        return ext

    def mode_basis_expansion(self, grid, S, regularization=0.0, Niter=1e4, rtol=1e-12, bad_modes=None,
                             RR_tol=1e-2, cond_tol=1e9, max_modes=None):
        """Basis.basis_expansion (cmbbest.pyx:320-490) with the contractions on the device.
        S: n_shapes x npts shape values at the tetrapyd points; max_modes None: every mode.
        Returns (alpha, shape_cov, corr, mse, selected)."""
        keep, args = self._modal(grid)
        S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, args[4])
        ns, n = S.shape[0], args[8]
        alpha, cov = np.zeros(ns * n), np.zeros(ns * ns)
        corr, mse = np.zeros(ns), np.zeros(ns)
        sel = np.zeros(ns * n, np.uint8)
        u8 = C.POINTER(C.c_uint8)
        bad = None if bad_modes is None else np.ascontiguousarray(bad_modes, dtype=np.uint8).reshape(n)
        badp = None if bad is None else bad.ctypes.data_as(u8)
        mm = 0 if max_modes is None else int(max_modes)
        _check(lib().aar_mode_basis_expansion(self.h, _dptr(alpha), _dptr(cov), _dptr(corr), _dptr(mse),
                                              sel.ctypes.data_as(u8), _dptr(S), ns, *args, badp,
                                              float(regularization), int(Niter), float(rtol), float(RR_tol),
                                              float(cond_tol), mm), "aar_mode_basis_expansion")
        return alpha.reshape(ns, n), cov.reshape(ns, ns), corr, mse, sel.reshape(ns, n).astype(bool)

    def ranlxd(self, seed, n):
        out = np.empty(n, np.float64)
        _check(lib().aar_ranlxd_device(self.h, seed, n, _dptr(out)), "aar_ranlxd_device")
        return out


class Job:
    def __init__(self, ctx, params, tabs, precision=AAR_FP64, schedule=AAR_SCHED_BATCHED):
        self.ctx, self.params, self.tabs = ctx, params, tabs
        h = C.c_void_p()
        _check(lib().aar_job_prepare(ctx.h, C.byref(params), C.byref(tabs.c), precision, schedule,
                                     C.byref(h)), "aar_job_prepare")
        self.h = h
        self.N = params.Nmesh
        self.npart = params.GlassTileFac ** 3 * tabs.nglass

    def run(self, stream=None):
        _check(lib().aar_job_run(self.h, stream), "aar_job_run")
        return self

    def reseed(self, seed, fnl=None):
        """Make the next run a new IC of this job's tables with `seed` (and `fnl`): the seed table
        is built now on the host, while earlier runs may still execute on the device."""
        fnl = self.params.Fnl if fnl is None else float(fnl)
        _check(lib().aar_job_reseed(self.h, int(seed), fnl), "aar_job_reseed")
        self.params.Seed, self.params.Fnl = int(seed), fnl
        return self

    def sync(self):
        _check(lib().aar_job_sync(self.h), "aar_job_sync")
        return lib().aar_job_last_seconds(self.h)

    def fetch(self):
        pos = np.empty(3 * self.npart, np.float32)
        vel = np.empty(3 * self.npart, np.float32)
        fpt = C.POINTER(C.c_float)
        _check(lib().aar_job_fetch(self.h, pos.ctypes.data_as(fpt), vel.ctypes.data_as(fpt)), "aar_job_fetch")
        return pos.reshape(-1, 3), vel.reshape(-1, 3)

    def field(self, name):
        N = self.N
        which = FIELDS[name]
        if which <= 1:
            out = np.empty(2 * N * N * (N // 2 + 1), np.float64)
        else:
            out = np.empty(N ** 3, np.float64)
        _check(lib().aar_job_fetch_field(self.h, which, _dptr(out)), "aar_job_fetch_field")
        if which <= 1:
            return out.view(np.complex128).reshape(N, N, N // 2 + 1)
        return out.reshape(N, N, N)

    def field_box(self, name, lo, shape):
        """Sub-box [lo, lo + shape) of a real field (displacements or potentials)."""
        out = np.empty(int(np.prod(shape)), np.float64)
        _check(lib().aar_job_fetch_field_box(self.h, FIELDS[name], *[int(v) for v in lo],
                                             *[int(v) for v in shape], _dptr(out)), "aar_job_fetch_field_box")
        return out.reshape(shape)

    def ng_statistics(self, edges):
        """Binned P(k) of phi_G and cross bispectrum <phi_G phi_G psi> with their expectations
        (aar_job_ng_statistics); bispectra indexed [b1, b2, b3]."""
        edges = np.ascontiguousarray(edges, np.float64)
        nb = len(edges) - 1
        r = {k: np.zeros(nb) for k in ("pk_meas", "pk_theory", "kmean", "nmodes")}
        r.update({k: np.zeros(nb ** 3) for k in ("b_meas", "b_theory", "ntri")})
        _check(lib().aar_job_ng_statistics(self.h, nb, _dptr(edges), *[_dptr(r[k]) for k in
                                           ("pk_meas", "pk_theory", "kmean", "nmodes", "b_meas", "b_theory",
                                            "ntri")]), "aar_job_ng_statistics")
        for k in ("b_meas", "b_theory", "ntri"):
            r[k] = r[k].reshape(nb, nb, nb)
        return r

    def set_graph(self, on=True):
        _check(lib().aar_job_set_graph(self.h, int(on)), "aar_job_set_graph")

    def jit(self):
        """1: the NG sum's fused z pass runs a kernel compiled for this job's schedule; 0: the
        library's; otherwise the status of a failed compile."""
        return lib().aar_job_jit(self.h)

    def set_profiling(self, on=True):
        _check(lib().aar_job_set_profiling(self.h, int(on)), "aar_job_set_profiling")

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(KINDS):
            n, s, b = C.c_int(0), C.c_double(0), C.c_double(0)
            _check(lib().aar_job_kernel_stats(self.h, k, C.byref(n), C.byref(s), C.byref(b)), "stats")
            if n.value:
                out[name] = dict(launches=n.value, seconds=s.value, bytes=b.value)
        return out

    def close(self):
        if self.h:
            lib().aar_job_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop on one device (aar_server_*).

    submit() returns a ticket at once; wait() blocks until that job completed
    and returns (latency_s, queue_s).  Jobs submitted with keep=True hold
    their results until release()."""

    def __init__(self, ctx, hbm_budget, max_active=4):
        h = C.c_void_p()
        _check(lib().aar_server_create(ctx.h, int(hbm_budget), int(max_active), C.byref(h)), "aar_server_create")
        self.h, self.ctx = h, ctx
        self.npart = {}

    def submit(self, params, tabs, precision=AAR_FP32, schedule=AAR_SCHED_BATCHED, keep=False):
        t = C.c_longlong(0)
        _check(lib().aar_server_submit(self.h, C.byref(params), C.byref(tabs.c), precision, schedule, int(keep),
                                       C.byref(t)), "aar_server_submit")
        self.npart[t.value] = params.GlassTileFac ** 3 * tabs.nglass
        return t.value

    def wait(self, ticket):
        lat, q = C.c_double(0), C.c_double(0)
        _check(lib().aar_server_wait(self.h, ticket, C.byref(lat), C.byref(q)), "aar_server_wait")
        return lat.value, q.value

    def fetch(self, ticket):
        n = self.npart[ticket]
        pos = np.empty(3 * n, np.float32)
        vel = np.empty(3 * n, np.float32)
        fpt = C.POINTER(C.c_float)
        _check(lib().aar_server_fetch(self.h, ticket, pos.ctypes.data_as(fpt), vel.ctypes.data_as(fpt)),
               "aar_server_fetch")
        return pos.reshape(-1, 3), vel.reshape(-1, 3)

    def release(self, ticket):
        _check(lib().aar_server_release(self.h, ticket), "aar_server_release")
        self.npart.pop(ticket, None)

    def stats(self):
        s, c, pb = C.c_longlong(0), C.c_longlong(0), C.c_longlong(0)
        pa = C.c_int(0)
        _check(lib().aar_server_stats(self.h, C.byref(s), C.byref(c), C.byref(pa), C.byref(pb)), "aar_server_stats")
        return dict(submitted=s.value, completed=c.value, peak_active=pa.value, peak_bytes=pb.value)

    def close(self):
        if self.h:
            lib().aar_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
