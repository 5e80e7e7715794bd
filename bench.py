#!/usr/bin/env python3
"""Throughput of the NG-2LPT initial-condition stage on one MI355X.

One step = one whole IC job with a fresh seed: the host seed table of that
seed (built while the previous job still runs on the device) and its upload,
the Gaussian potential, the modal non-Gaussian sum, 2LPT, and particle
positions and velocities resident in HBM.  The job's tables (modal schedule,
k-space and transfer tables, glass) are uploaded when the job is first
prepared; that cold first job is timed on its own.  Headline workload:
Nmesh = Npart = 512^3, Lbox = 2000 Mpc/h, the ScalarI (mass 1) template
decomposed on the device into 16 basis modes, X0 from its closed form at
Nmax = 512, Eisenstein-Hu transfer, fNL = 100, FP32 batched schedule.  The
FP64 pairwise schedule, the 1024^3 local+equilateral job, the serving stream,
the decomposition and the CPU oracle are reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


TEMPLATES = {
    "scalari16": "ScalarI mass=1 decomposed on the device into 16 basis modes (ModeTol 1e-4, MaxModeCount 150)",
    "scalari": "ScalarI mass=1 15-mode table (Aarambam/defaults/AlphaTable.dat)",
    "le": "local+equilateral closed-form table, 12 modes",
    "local": "local closed-form table, 8 modes",
    "equil": "equilateral closed-form table, 12 modes",
    "dense12": "synthetic dense 12-mode table (60 triples, alpha ~ N(0,1), 39 product pairs)",
}

# the reference's shipped 15-mode ScalarI table (Aarambam/defaults/AlphaTable.dat), a data file
SHIPPED_ALPHA = os.path.join(ROOT, "tests", "golden", "AlphaTable.dat")

_SCALARI16 = {}


def scalari16_alpha(ctx, ns, box, nmax):
    """AlphaTable of the ScalarI (mass 1) template for 16 basis modes, made by
    aarambam_amd.decompose the way Aarambam/basis/decompose.py:84-174 makes it (150-point k grid
    from 2 pi / L to 2 pi Nmax / L, the shape symmetrised over the orderings, basis_expansion with
    rtol 1e-10, 1000 CG iterations, its mode mask, ModeTol 1e-4, MaxModeCount 150), the expansion
    on the device.  Cached per (n_s, box, Nmax)."""
    from aarambam_amd import decompose
    from aarambam_amd import workloads as W
    key = (ns, box, nmax)
    if key not in _SCALARI16:
        t0 = time.perf_counter()
        r = decompose.decompose(ctx, lambda a, b, c: W.scalar_i(a, b, c, ns, 1.0), 16, ns, box, nmax, 1e-4, 150)
        _SCALARI16[key] = (W.alpha_table(16, r["AlphaTable"]), r["correlation"],
                           int((r["AlphaTable"][:, 3] != 0).sum()), time.perf_counter() - t0)
    return _SCALARI16[key]


def workload(N=512, box=2000.0, template="scalari16", seed=1234, fnl=100.0, ctx=None):
    from aarambam_amd import workloads as W
    nm = {"scalari16": 16, "scalari": 15, "le": 12, "local": 8, "equil": 12, "dense12": 12}[template]
    cfg = W.config(Nmesh=N, Nsample=N, Nmax=N, N_modes=nm, Box=box, GlassTileFac=max(1, N // 64), Fnl=fnl,
                   Seed=seed)
    x0 = W.x0_table(nm, cfg["PrimordialIndex"], box, N)
    if template == "scalari16":
        alpha = scalari16_alpha(ctx, cfg["PrimordialIndex"], box, N)[0]
    elif template == "scalari":
        alpha = load_alpha(SHIPPED_ALPHA, 15)
    else:
        rows = {"local": W.local_rows, "equil": W.equilateral_rows, "le": W.le_rows,
                "dense12": W.dense12_rows}[template]()
        alpha = W.alpha_table(nm, rows)
    glass = W.lattice_glass(64 if N >= 64 else N, 100.0)
    return cfg, alpha, x0, glass, 100.0


def load_alpha(path, nm):
    import ctypes as C

    import aarambam_amd as aar
    a = np.empty(nm ** 3)
    rc = aar.lib().aar_load_alpha(path.encode(), nm, a.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise RuntimeError(f"aar_load_alpha failed: {rc}")
    return a.reshape(nm, nm, nm)


def modal_sizes(alpha):
    """(M, K): distinct mode fields and distinct outer indices of the regrouped sum."""
    nm = alpha.shape[0]
    modes, outer = set(), set()
    for i in range(nm):
        for j in range(i, nm):
            for k in range(j, nm):
                if alpha[i, j, k] != 0:
                    modes |= {i, j}
                    outer.add(k)
    if alpha[0, 1, 2] != 0 and alpha[1, 1, 1] != 0:  # the two divergence-cancelling terms
        modes |= {0, 1, 2}
        outer |= {1, 2}
    return len(modes), len(outer)


def survey_bytes(N, M, K, b):
    """Algorithmic bytes of one job: S_b (94 + 7M + 8K) + 24 N^3 (one spectrum S_b = N^2 (N/2+1) 2 b)."""
    S = N * N * (N // 2 + 1) * 2 * b
    return S * (94 + 7 * M + 8 * K) + 24 * N ** 3


ERR_FIELDS = ["disp_x", "disp_y", "disp_z", "disp2_x", "disp2_y", "disp2_z", "phing_real"]
ERR_SEED = 4321  # both precisions re-run this seed after their timed regions for the error figure


def time_stream(aar, ctx, cfg, tabs, prec, sched, steps, warmup, profile, graph=False, box=0, dev_steps=0):
    """A stream of jobs of one configuration with fresh seeds.  cold: prepare (cosmology, tables
    uploaded once, seed table) + first run + sync.  Then `warmup` and `steps` jobs, each a host seed
    table for a new seed (aar_job_reseed, overlapping the previous job on the device), its upload and
    the whole job; the steps are timed from the first reseed to the last job's completion.
    dev_steps: afterwards, that many runs of one seed, each waited for: the device-pipeline time."""
    t0 = time.perf_counter()
    job = aar.Job(ctx, aar.params_from_dict(cfg), tabs, prec, sched)
    job.set_graph(graph)
    job.run()
    job.sync()
    cold = time.perf_counter() - t0
    seed0 = int(cfg["Seed"])
    for w in range(warmup):
        job.reseed(seed0 + 1 + w)
        job.run()
    job.sync()
    if profile:
        job.set_profiling(True)  # per-kind HIP events around every launch of the timed jobs
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        job.reseed(seed0 + 1 + warmup + k)
        job.run()
    job.sync()
    ctx.sync()
    wall = time.perf_counter() - t0
    stats = job.kernel_stats() if profile else {}
    job.set_profiling(False)
    dev = []
    for _ in range(dev_steps):
        job.run()
        dev.append(job.sync())
    info = dict(wall=wall, cold=cold, stats=stats, device=float(np.mean(dev)) if dev else None)
    if box:  # a box^3 corner of the displacement and potential fields of one seed
        job.reseed(ERR_SEED)
        job.run()
        job.sync()
        info["box"] = {n: job.field_box(n, (0, 0, 0), (box, box, box)) for n in ERR_FIELDS}
    job.close()
    return info


def rel_errors(a, b):
    err = {n: float(np.linalg.norm(a[n] - b[n]) / np.linalg.norm(b[n])) for n in ERR_FIELDS}
    return {"max": float(f"{max(err.values()):.3g}"), "fields": {n: float(f"{v:.3g}") for n, v in err.items()}}


def kind_gbs(stats):
    """Algorithmic GB/s of every kernel kind: its summed bytes over its summed HIP-event time."""
    return {k: round(v["bytes"] / v["seconds"] / 1e9, 1) for k, v in stats.items() if v["seconds"] > 0}


def roofline_of(stats):
    dom = max(stats, key=lambda k: stats[k]["seconds"])
    d = stats[dom]
    achieved = d["bytes"] / d["seconds"] / 1e9
    return dom, d, achieved


def north_star_1024(aar, ctx, steps=3, fp64_steps=2):
    """1024^3, Lbox 2000 Mpc/h, local + equilateral (12 modes), fNL 100: the FP32 batched stream
    against the FP64 pairwise stream, the job roofline, and the FP32 error on a 256^3 sub-box."""
    N = 1024
    cfg, alpha, x0, glass, gbox = workload(N, 2000.0, "le")
    tabs = aar.tables(alpha, x0, glass, gbox)
    psec, plan = ctx.autotune_plans(N, aar.AAR_FP32, nf=8)
    r32 = time_stream(aar, ctx, cfg, tabs, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, steps, 1, True, box=256)
    r64 = time_stream(aar, ctx, cfg, tabs, aar.AAR_FP64, aar.AAR_SCHED_PLAIN, fp64_steps, 0, False, box=256)
    M, K = modal_sizes(alpha)
    ms32, ms64 = r32["wall"] / steps * 1e3, r64["wall"] / fp64_steps * 1e3
    dom, d, achieved = roofline_of(r32["stats"])
    B = survey_bytes(N, M, K, 4)
    return {"workload": "NG-2LPT IC, Nmesh=Npart=1024^3, Lbox=2000 Mpc/h, local+equilateral closed-form table "
                        f"(12 modes, M={M}, K={K}), EH P(k), fNL=100, fresh seed per job",
            "fp32_batched": {"ms_per_job": round(ms32, 2), "mcells_s": round(N ** 3 / ms32 / 1e3, 1),
                             "cold_job_ms": round(r32["cold"] * 1e3, 1), "steps": steps},
            "fp64_plain": {"ms_per_job": round(ms64, 2), "mcells_s": round(N ** 3 / ms64 / 1e3, 1),
                           "steps": fp64_steps},
            "fp32_over_fp64": round(ms64 / ms32, 3),
            "autotune": {"radix_plan": plan, "us_per_xy_pair_8_fields": [round(v * 1e6, 1) for v in psec],
                         "xy_tile_variant": [aar.lib().aar_xy_tile_variant(N, aar.AAR_FP32, a) for a in (0, 1, 2)],
                         "xy_pair_group_cols": aar.lib().aar_xy_group_cols(N, aar.AAR_FP32),
                         "xy_split_group_cols": aar.lib().aar_xy_split_group_cols(N, aar.AAR_FP32)},
            "job_roofline": {"bytes_per_job": B, "frac": round(B / (ms32 * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)},
            "xy_access_pattern_gbs": xy_pattern_gbs(ctx, N, nf=4),
            "dominant_kernel": {"kernel": dom, "achieved_gbs": round(achieved, 1),
                                "frac": round(achieved / HBM_PEAK_GBS, 4),
                                "avg_launch_us": round(d["seconds"] / d["launches"] * 1e6, 2)},
            "kernel_gbs": kind_gbs(r32["stats"]),
            "kernel_time_share": {k: round(v["seconds"] / sum(x["seconds"] for x in r32["stats"].values()), 4)
                                  for k, v in r32["stats"].items()},
            "fp32_rel_l2_vs_fp64": dict(sub_box=256, **rel_errors(r32["box"], r64["box"]))}


def serving(aar, ctx, n_jobs=64, load=0.75, budget_gib=96, max_active=4, seed=2024):
    """Poisson stream of IC jobs through the continuous-batching server.

    Nmesh uniform over {128, 256, 512}, template uniform over {local,
    equilateral, ScalarI mass=1}, seeds and fNL ~ U[-200, 200] from a fixed-seed
    PCG64.  The arrival rate is `load` times the capacity measured first: the
    inverse of the mean time of one job of each (Nmesh, template) run alone.
    Latency = completion - submission of each job."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes, tpls = (128, 256, 512), ("local", "equil", "scalari")
    cache = {}

    def inputs(N, tpl, sd, fnl):
        cfg, alpha, x0, glass, gbox = workload(N, 2000.0 * N / 512, tpl, sd, fnl)
        key = (N, tpl)
        if key not in cache:
            cache[key] = aar.tables(alpha, x0, glass, gbox)
        return aar.params_from_dict(cfg), cache[key]

    srv = aar.Server(ctx, int(budget_gib) << 30, max_active)
    alone = {}
    for N in sizes:
        for tpl in tpls:
            p, t = inputs(N, tpl, 1, 100.0)
            for rep in range(2):  # the first run of a shape warms its code paths
                tk = srv.submit(p, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
                lat, _ = srv.wait(tk)
                srv.release(tk)
            alone[f"{N}/{tpl}"] = lat
    mean_service = float(np.mean(list(alone.values())))
    rate = load / mean_service
    specs = [(int(rng.choice(sizes)), str(rng.choice(tpls)), int(rng.integers(1, 2 ** 31 - 1)),
              float(rng.uniform(-200, 200))) for _ in range(n_jobs)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    jobs = [inputs(*sp) for sp in specs]
    tickets = []
    t0 = time.perf_counter()
    for (p, t), ta in zip(jobs, arrivals):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        tickets.append(srv.submit(p, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED))
    lats = []
    for tk in tickets:
        lats.append(srv.wait(tk)[0])
        srv.release(tk)
    span = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lats = np.array(lats) * 1e3
    cells = sum(sp[0] ** 3 for sp in specs)
    return {"jobs": n_jobs, "p50_ms": round(float(np.percentile(lats, 50)), 3),
            "p99_ms": round(float(np.percentile(lats, 99)), 3), "mean_ms": round(float(lats.mean()), 3),
            "alone_ms": {k: round(v * 1e3, 3) for k, v in alone.items()},
            "arrival_rate_per_s": round(rate, 2), "load": load,
            "throughput_mcells_s": round(cells / span / 1e6, 1), "span_s": round(span, 3),
            "max_active": max_active, "peak_active": st["peak_active"], "hbm_budget_gib": budget_gib,
            "peak_arena_gib": round(st["peak_bytes"] / 2 ** 30, 2),
            "mix": "Nmesh U{128,256,512}, template U{local, equilateral, ScalarI}, seed and fNL U[-200,200] "
                   "from PCG64(2024), FP32 batched, Lbox = 2000 Mpc/h * Nmesh/512"}


def decomposition(aar, ctx, p_max=16, nk=150, reps=3):
    """The decomposition's mode-bispectrum covariance (Aarambam/basis/tetrapyd.c:111-208) on the
    FP64 matrix cores, at the Basis sizes of the reference's decompose.py (k_grid_size 150) with 16
    modes (816 3-D modes); the reference's own C (oracle/_ref) timed beside it on a 40-point grid."""
    from aarambam_amd import decompose
    from oracle import pyoracle as orc

    def grid(nk_):  # the Basis grid of the reference's decomposition defaults (L 1000, Nmax 256)
        return decompose.basis_grid(p_max, 0.9660499, 2 * np.pi / 1000, 2 * np.pi / 1000 * 256, nk_)

    g = grid(nk)
    n, T = len(g["p1"]), len(g["w"])
    ctx.mode_covariance(g)  # warm
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.mode_covariance(g)
        ts.append(time.perf_counter() - t0)
    dt = min(ts)
    flops = float(n) * (n + 1) * T  # the symmetric matrix's distinct entries, one multiply-add per point
    out = {"modes_3d": n, "tetrapyd_points": T, "ms": round(dt * 1e3, 2),
           "useful_tflops": round(flops / dt / 1e12, 2), "fp64_mfma_peak_tflops": 78.6,
           "note": "wall time of aar_mode_bispectra_covariance incl. input upload and result copy"}
    small = grid(40)
    t0 = time.perf_counter()
    ctx.mode_covariance(small)
    gdt = time.perf_counter() - t0
    lib = orc.ref_lib()
    t0 = time.perf_counter()
    orc.mode_covariance(small, use_ref=lib is not None)
    cdt = time.perf_counter() - t0
    out["cpu_sample"] = {"kind": "reference" if lib is not None else "port", "k_grid": 40,
                         "points": len(small["w"]), "cpu_s": round(cdt, 3), "gpu_s": round(gdt, 4),
                         "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    out["expansion"] = expansion(ctx)
    return out


def expansion(ctx):
    """Basis.basis_expansion (Aarambam/basis/cmbbest.pyx:320-490) of the local shape at the settings
    of the reference's decompose.py and MakeICs example (N_modes 15, k_grid_size 150, ModeTol 1e-5,
    MaxModeCount 150), and with a tolerance that runs the selection loop to 150 modes; the oracle's
    numpy restatement times a 10-mode, 60-point-grid sample."""
    from aarambam_amd import decompose
    from oracle import pymodal

    def basis_grid(p_max, k_grid_size):
        return decompose.basis_grid(p_max, 0.9660499, 2 * np.pi / 1000, 2 * np.pi / 1000 * 256, k_grid_size)

    def local(g):
        k = g["k"]
        k1, k2, k3 = k[g["i1"]], k[g["i2"]], k[g["i3"]]
        return ((k1 * k1 / (k2 * k3) + k2 * k2 / (k1 * k3) + k3 * k3 / (k1 * k2)) / 3)[None, :]

    g = basis_grid(15, 150)
    S = local(g)
    res = {"modes_3d": len(g["p1"]), "tetrapyd_points": len(g["w"])}
    for name, tol in [("ModeTol_1e-5", 1e-5), ("ModeTol_1e-12", 1e-12)]:
        t0 = time.perf_counter()
        r = ctx.mode_basis_expansion(g, S, rtol=1e-10, Niter=1e3, RR_tol=tol, max_modes=150)
        res[name] = {"s": round(time.perf_counter() - t0, 3), "selected": int(r[4][0].sum()),
                     "correlation": float(r[2][0])}
    small = basis_grid(10, 60)
    Ss = local(small)
    kw = dict(rtol=1e-10, Niter=1e3, RR_tol=1e-5, max_modes=150)
    t0 = time.perf_counter()
    rs = ctx.mode_basis_expansion(small, Ss, **kw)
    gdt = time.perf_counter() - t0
    t0 = time.perf_counter()
    ro = pymodal.basis_expansion(small, Ss, **kw)
    cdt = time.perf_counter() - t0
    res["cpu_sample"] = {"kind": "port", "modes_3d": len(small["p1"]), "points": len(small["w"]),
                         "cpu_s": round(cdt, 3), "gpu_s": round(gdt, 4),
                         "same_selection": bool((ro[4] == rs[4]).all())}
    return res


def copy_peak(ctx, N=512, nf=8, iters=10):
    """HBM bandwidth of a plain streaming copy kernel (reads + writes nf spectra)."""
    sec = ctx.bench_pass(N, 4, nf, iters)
    S = N * N * (N // 2 + 1) * 8
    return 2 * nf * S / sec / 1e9


def xy_pattern_gbs(ctx, N, nf=8, iters=5):
    """Rate of the x and y passes' tile access with no transform (aar_bench_pass kinds 5, 6):
    reads and writes of the kz < N/2 block of nf FP32 fields."""
    nbytes = 2 * nf * N * N * (N // 2) * 8
    return {ax: round(nbytes / ctx.bench_pass(N, k, nf, iters) / 1e9, 1) for ax, k in (("x", 5), ("y", 6))}


def profiled_traffic(kernel, N, template):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (counters scaled by the same counters on a copy kernel of known size); None when not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(f) >= 4 and f[0] == kernel and f[1] == str(N) and f[2] == template:
            try:
                return float(f[3])
            except ValueError:
                return None
    return None


def cpu_baseline(N=256, ctx=None, alpha=None, headline_n=512):
    """The oracle on one N^3 job of the timed workload's table, and its rate carried to the
    headline mesh: the oracle's work per cell grows as the FFTs' log2(N^3) (its other loops are
    linear in the cells), so the rate at headline_n is the measured rate times
    log2(N^3) / log2(headline_n^3)."""
    from aarambam_amd import workloads as W
    from oracle import pyoracle as orc
    cfg, _, _, glass, gbox = workload(N, 2000.0 * N / 512, "le", ctx=ctx)
    nm = alpha.shape[0]
    cfg["N_modes"] = nm
    x0 = W.x0_table(nm, cfg["PrimordialIndex"], cfg["Box"], N)
    t0 = time.perf_counter()
    orc.run(cfg, alpha, x0, glass, gbox, want_fields=False)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    v = N ** 3 / dt / 1e6
    scale = np.log2(float(N) ** 3) / np.log2(float(headline_n) ** 3)
    return dict(value=round(v, 4), unit="Mcells/s", cores=cores, kind="port",
                sample=f"CPU oracle (C/OpenMP FP64, the reference's pairwise loops, long double accumulator), "
                       f"one {N}^3 job of the headline's modal table with Lbox scaled to {cfg['Box']:g}, "
                       f"{dt:.2f} s on {cores} threads",
                extrapolated_to_headline={"nmesh": headline_n, "value": round(v * scale, 4),
                                          "model": "rate x log2(N^3) / log2(512^3) (FFT work per cell)"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nmesh", type=int, default=512)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--template", default="scalari16", choices=sorted(TEMPLATES))
    ap.add_argument("--box", type=float, default=2000.0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-max-active", type=int, default=4, help="jobs the server keeps in flight at most")
    ap.add_argument("--no-decomp", action="store_true")
    ap.add_argument("--no-1024", action="store_true", help="skip the 1024^3 local+equilateral leg")
    ap.add_argument("--xy-group", type=int, default=-1,
                    help="pin the kz column groups of the x/y pass pairs (default: the autotuner's choice); "
                         "counter runs serialize every launch, which skews the autotuner, so they pin it")
    ap.add_argument("--xy-plan", type=int, default=-1,
                    help="pin the x/y radix plan (default: the autotuner's choice); counter runs pin it too")
    ap.add_argument("--modes-x", type=int, default=1, choices=(0, 1, 2),
                    help="filter the mode fields in the loads of their x pass: 1 when the filters are short "
                         "(default), 2 always, 0 never (written in k-space first)")
    ap.add_argument("--xy-split-group", type=int, default=-2,
                    help="pin the kz column groups of the multi-output x/y pass pairs (-1: the pairs' choice; "
                         "default: the autotuner's choice)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported (one IC job runs on one "
              "GPU, and independent jobs shard across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)
    import aarambam_amd as aar
    N = args.nmesh
    ctx = aar.Context(0)
    ctx.set_modes_x(args.modes_x)
    cfg, alpha, x0, glass, gbox = workload(N, args.box, args.template, ctx=ctx)
    tabs = aar.tables(alpha, x0, glass, gbox)
    # x / y tile shapes and x / y pair column groups for this mesh, picked by the autotuner
    # before the timed region
    pow2 = N & (N - 1) == 0  # other mesh sizes run the mixed-radix passes, which have no tuned shapes
    tsec, tvar, psec, pchosen, ssec, gsec = ([[0] * 4] * 2, [0, 0, 0], [0] * 3, 0, [0] * 9, [[0] * 4] * 3)
    if pow2:
        tsec, gsec, ssec = ctx.autotune_report(N, aar.AAR_FP32, nf=8)
        psec = [min([v for v in row if v > 0], default=-1.0) for row in gsec]
        tvar = [aar.lib().aar_xy_tile_variant(N, aar.AAR_FP32, a) for a in (0, 1, 2)]
        pchosen = aar.lib().aar_xy_radix_plan(N, aar.AAR_FP32)
    if args.xy_group >= 0 and aar.lib().aar_set_xy_group_cols(N, aar.AAR_FP32, args.xy_group) != 0:
        sys.exit(f"--xy-group {args.xy_group}: not a valid column group for N = {N}")
    if args.xy_plan >= 0 and aar.lib().aar_set_xy_radix_plan(N, aar.AAR_FP32, args.xy_plan) != 0:
        sys.exit(f"--xy-plan {args.xy_plan}: not a valid radix plan for N = {N}")
    if args.xy_split_group >= -1 and aar.lib().aar_set_xy_split_group_cols(N, aar.AAR_FP32, args.xy_split_group) != 0:
        sys.exit(f"--xy-split-group {args.xy_split_group}: not a valid column group for N = {N}")
    xy_group = aar.lib().aar_xy_group_cols(N, aar.AAR_FP32)
    xy_split_group = aar.lib().aar_xy_split_group_cols(N, aar.AAR_FP32)
    # timed region: a stream of jobs with fresh seeds, eager launches each bracketed by HIP events
    eb = 0 if args.no_fp64 else min(128, N)
    r32 = time_stream(aar, ctx, cfg, tabs, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, args.steps, args.warmup, True,
                      box=eb, dev_steps=args.steps)
    rg = None
    if not args.no_graph:  # the same stream replaying one captured hipGraph per job
        rg = time_stream(aar, ctx, cfg, tabs, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, args.steps, args.warmup, False,
                         graph=True)
    ms = r32["wall"] / args.steps * 1e3
    value = N ** 3 * args.steps / r32["wall"] / 1e6
    stats = r32["stats"]
    dom, d, achieved = roofline_of(stats)
    M, K = modal_sizes(alpha)
    decomp_note = ""
    if args.template == "scalari16":
        _, corr16, nz16, dt16 = scalari16_alpha(ctx, cfg["PrimordialIndex"], args.box, N)
        decomp_note = f": {nz16} nonzero coefficients, correlation {corr16:.6f}"
    ktime = sum(v["seconds"] for v in stats.values())
    out = {
        "metric": "Mcells/s for NG-2LPT IC at Nmesh=512, FP32 vs FP64; p50/p99 job latency",
        "value": round(value, 3), "unit": "Mcells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Gaussian field, lattice glass)",
        "config": {"workload": f"NG-2LPT IC, Nmesh=Npart={N}^3, Lbox={args.box:g} Mpc/h, "
                               f"{TEMPLATES[args.template]}{decomp_note} (M={M}, K={K}), EH P(k), fNL=100, "
                               f"a fresh seed per job (1235, 1236, ...)",
                   "precision": "fp32", "schedule": "batched",
                   "step": "host seed table of the job's seed (overlapping the previous job) + its upload + "
                           "the whole job on the device; tables resident from the first (cold) job"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(dom, N, args.template),
                     "algorithmic_bytes_per_launch": round(d["bytes"] / d["launches"]),
                     "avg_launch_us": round(d["seconds"] / d["launches"] * 1e6, 2)},
        "kernel_time_share": {k: round(v["seconds"] / ktime, 4) for k, v in stats.items()},
        "kernel_gbs": kind_gbs(stats),
        "job_roofline": {"bytes_per_job": survey_bytes(N, M, K, 4),
                         "frac": round(survey_bytes(N, M, K, 4) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)},
        "device_pipeline": {"ms_per_job": round(r32["device"] * 1e3, 3),
                            "mcells_s": round(N ** 3 / r32["device"] / 1e6, 1),
                            "note": "one seed, each job waited for: device time between the job's first and "
                                    "last kernel (HIP events), no host seed table"},
        "cold_job_ms": round(r32["cold"] * 1e3, 2),
    }
    if args.template == "scalari16":
        out["table_decomposition_s"] = round(dt16, 3)  # the 16-mode ScalarI table, made on the device first
    if rg is not None:
        out["hipgraph"] = {"ms_per_step": round(rg["wall"] / args.steps * 1e3, 3),
                           "value": round(N ** 3 * args.steps / rg["wall"] / 1e6, 3)}
    if not args.no_fp64:
        r64 = time_stream(aar, ctx, cfg, tabs, aar.AAR_FP64, aar.AAR_SCHED_PLAIN, args.fp64_steps, 1, False, box=eb)
        v64 = N ** 3 * args.fp64_steps / r64["wall"] / 1e6
        out["fp64_plain"] = {"value": round(v64, 3), "unit": "Mcells/s",
                             "ms_per_step": round(r64["wall"] / args.fp64_steps * 1e3, 3),
                             "fp32_over_fp64": round(value / v64, 3)}
        # accuracy of the timed FP32 path: relative L2 error against the FP64 pairwise path, same seed,
        # on a box^3 corner of every displacement field and of the non-Gaussian potential
        out["fp32_rel_l2_vs_fp64"] = dict(sub_box=eb, seed=ERR_SEED, **rel_errors(r32["box"], r64["box"]))
    out["autotune"] = None if not pow2 else {"xy_tile_variant": {"x": tvar[0], "y": tvar[1], "y_split": tvar[2]},
                       "us_per_launch_8_fields": {ax: [round(v * 1e6, 1) if v > 0 else None for v in tsec[i]]
                                                  for i, ax in enumerate(("x", "y"))},
                       "variants": "0: 128-B tile columns, 1: 64-B, 2: 256-B, 3: 128-B, four tiles per workgroup with the next tile's loads in flight (same radix plan)",
                       "xy_pair_group_cols": xy_group,
                       "us_per_xy_pair_8_fields_by_group": {
                           "groups": "whole passes, all columns, halves, quarters",
                           "per_plan": [[round(v * 1e6, 1) if v > 0 else None for v in row] for row in gsec]},
                       "xy_split_group_cols": {"chosen": xy_split_group,
                                               "us_per_call_9_outputs": [round(v * 1e6, 1) if v > 0 else None
                                                                         for v in ssec[:5]],
                                               "groups": "whole passes, all columns, halves, quarters, eighths"},
                       "xy_split_y_tile_variant": {"chosen": tvar[2],
                                                   "us_per_call_9_outputs": [round(v * 1e6, 1) if v > 0 else None
                                                                             for v in ssec[5:]]},
                       "radix_plan": {"chosen": pchosen,
                                      "us_per_xy_pair_8_fields": [round(v * 1e6, 1) if v > 0 else None for v in psec],
                                      "plans": "0: radices spread evenly (512 = 8 8 8, 1024 = 16 8 8), "
                                               "1: 16 first (16 16 2, 16 16 4), 2: 16 last (2 16 16, 4 16 16)"}}
    out["copy_peak_gbs"] = round(copy_peak(ctx, 512), 1)
    if pow2:
        out["xy_access_pattern_gbs"] = xy_pattern_gbs(ctx, N)
    if not args.no_1024:
        out["north_star_1024"] = north_star_1024(aar, ctx)
    if not args.no_serve:
        out["serving"] = serving(aar, ctx, args.serve_jobs, max_active=args.serve_max_active)
    if not args.no_decomp:
        out["decomposition"] = decomposition(aar, ctx)
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, alpha=alpha, headline_n=N)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
