#!/usr/bin/env python3
"""Throughput of the NG-2LPT initial-condition stage on one MI355X.

One step = one whole IC job on the device (Gaussian potential from the seed
table, modal non-Gaussian sum, 2LPT, particle positions and velocities), with
all inputs resident in HBM.  Workload: Nmesh = Npart = 512^3, Lbox = 2000
Mpc/h, the reference's shipped ScalarI (mass 1) modal table
(Aarambam/defaults/AlphaTable.dat, 15 modes) with X0 from its closed form at
Nmax = 512, Eisenstein-Hu transfer, fNL = 100.  Headline: FP32 batched
schedule; the FP64 pairwise schedule and the CPU oracle are reported beside it.
"""
import argparse
import itertools
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


_SCALARI16 = {}


def scalari16_alpha(ctx, ns, box, nmax):
    """AlphaTable of the ScalarI (mass 1) template for 16 basis modes, made by
    aarambam_amd.decompose the way Aarambam/basis/decompose.py:84-174 makes it (150-point k grid
    from 2 pi / L to 2 pi Nmax / L, the shape symmetrised over the orderings, basis_expansion with
    rtol 1e-10, 1000 CG iterations, its mode mask, ModeTol 1e-4, MaxModeCount 150), the expansion
    on the device.  Cached per (n_s, box, Nmax)."""
    from aarambam_amd import decompose
    from tests import templates
    key = (ns, box, nmax)
    if key not in _SCALARI16:
        r = decompose.decompose(ctx, lambda a, b, c: templates.scalar_i(a, b, c, ns, 1.0), 16, ns, box, nmax,
                                1e-4, 150)
        table = np.zeros((16, 16, 16))
        for n1, n2, n3, a in r["AlphaTable"]:  # main.c:231-241 fills all six orders
            for q in itertools.permutations((int(n1), int(n2), int(n3))):
                table[q] = a
        _SCALARI16[key] = (table, r["correlation"], int((r["AlphaTable"][:, 3] != 0).sum()))
    return _SCALARI16[key]


def x0_table(nm, ns, box, nmax):
    from tests import cases
    return cases.x0_closed_form(nm, ns, box, nmax)


def workload(N=512, box=2000.0, template="scalari16", seed=1234, fnl=100.0, ctx=None):
    from tests import cases
    nm = {"scalari16": 16, "scalari": 15, "le": 12, "local": 8, "equil": 12, "dense12": 12}[template]
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=nm, Box=box, GlassTileFac=max(1, N // 64),
                            Fnl=fnl, Seed=seed)
    x0 = cases.x0_closed_form(nm, cfg["PrimordialIndex"], box, N)
    if template == "scalari16":
        alpha = scalari16_alpha(ctx, cfg["PrimordialIndex"], box, N)[0]
    elif template == "scalari":
        alpha = load_alpha(os.path.join(cases.GOLDEN, "AlphaTable.dat"), 15)
    else:
        rows = {"local": cases.local_rows, "equil": cases.equilateral_rows, "le": cases.le_rows,
                "dense12": cases.dense12_rows}[template]()
        alpha = np.zeros((nm, nm, nm))
        for n1, n2, n3, a in rows:  # main.c:231-241 fills all six permutations
            n1, n2, n3 = int(n1), int(n2), int(n3)
            for p in [(n1, n2, n3), (n1, n3, n2), (n2, n1, n3), (n2, n3, n1), (n3, n1, n2), (n3, n2, n1)]:
                alpha[p] = a
    glass = cases.lattice_glass(64 if N >= 64 else N, 100.0)
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


def time_job(aar, ctx, cfg, alpha, x0, glass, gbox, prec, sched, steps, warmup, profile, graph=False, box=0):
    t = aar.tables(alpha, x0, glass, gbox)
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.set_graph(graph)
    for _ in range(warmup):
        job.run()
    job.sync()
    if profile:
        job.set_profiling(True)  # resets the counters, keeps the captured graph
    ctx.sync()
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(steps):
        job.run()
        dev += job.sync()
    ctx.sync()
    wall = time.perf_counter() - t0
    stats = job.kernel_stats() if profile else {}
    info = dict(wall=wall, device=dev, stats=stats, M=None)
    if box:  # a box^3 corner of the displacement and potential fields, after the timed region
        info["box"] = {n: job.field_box(n, (0, 0, 0), (box, box, box)) for n in ERR_FIELDS}
    job.close()
    return info


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
    from oracle import pyoracle as orc
    from tests import modal
    g = modal.basis_grid(p_max=p_max, k_grid_size=nk)
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
    small = modal.basis_grid(p_max=p_max, k_grid_size=40)
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
    from oracle import pymodal
    from tests import modal

    def local(g):
        k = g["k"]
        k1, k2, k3 = k[g["i1"]], k[g["i2"]], k[g["i3"]]
        return ((k1 * k1 / (k2 * k3) + k2 * k2 / (k1 * k3) + k3 * k3 / (k1 * k2)) / 3)[None, :]

    g = modal.basis_grid(p_max=15, k_grid_size=150, n_s=0.9660499)
    S = local(g)
    res = {"modes_3d": len(g["p1"]), "tetrapyd_points": len(g["w"])}
    for name, tol in [("ModeTol_1e-5", 1e-5), ("ModeTol_1e-12", 1e-12)]:
        t0 = time.perf_counter()
        r = ctx.mode_basis_expansion(g, S, rtol=1e-10, Niter=1e3, RR_tol=tol, max_modes=150)
        res[name] = {"s": round(time.perf_counter() - t0, 3), "selected": int(r[4][0].sum()),
                     "correlation": float(r[2][0])}
    small = modal.basis_grid(p_max=10, k_grid_size=60, n_s=0.9660499)
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


def cpu_baseline(N=128, template="scalari16", ctx=None, alpha=None):
    """The oracle on one N^3 job of the timed workload's table (alpha, when given, is that table)."""
    from oracle import pyoracle as orc
    cfg, a, x0, glass, gbox = workload(N, 2000.0 * N / 512, template if alpha is None else "le", ctx=ctx)
    if alpha is not None:
        nm = alpha.shape[0]
        cfg["N_modes"] = nm
        x0 = x0_table(nm, cfg["PrimordialIndex"], cfg["Box"], N)
    else:
        alpha = a
    t0 = time.perf_counter()
    orc.run(cfg, alpha, x0, glass, gbox, want_fields=False)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=N ** 3 / dt / 1e6, unit="Mcells/s", cores=cores, kind="port",
                sample=f"CPU oracle (C/OpenMP FP64, reference's pairwise loops), one {N}^3 job of the "
                       f"same table, {dt:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
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
    ap.add_argument("--no-decomp", action="store_true")
    ap.add_argument("--xy-group", type=int, default=-1,
                    help="pin the kz column groups of the x/y pass pairs (default: the autotuner's choice); "
                         "counter runs serialize every launch, which skews the autotuner, so they pin it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported (one IC job runs on one "
              "GPU, and independent jobs shard across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)
    import aarambam_amd as aar
    N = args.nmesh
    ctx = aar.Context(0)
    cfg, alpha, x0, glass, gbox = workload(N, args.box, args.template, ctx=ctx)
    # x / y tile shapes and x / y pair column groups for this mesh, picked by the autotuner
    # before the timed region
    tsec, tvar = ctx.autotune(N, aar.AAR_FP32, nf=8)
    if args.xy_group >= 0 and aar.lib().aar_set_xy_group_cols(N, aar.AAR_FP32, args.xy_group) != 0:
        sys.exit(f"--xy-group {args.xy_group}: not a valid column group for N = {N}")
    xy_group = aar.lib().aar_xy_group_cols(N, aar.AAR_FP32)
    # timed region: eager launches, each bracketed by HIP events (per-kernel roofline)
    eb = 0 if args.no_fp64 else min(128, N)
    r32 = time_job(aar, ctx, cfg, alpha, x0, glass, gbox, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, args.steps,
                   args.warmup, True, box=eb)
    rg = None
    if not args.no_graph:  # the same job replayed as one captured hipGraph
        rg = time_job(aar, ctx, cfg, alpha, x0, glass, gbox, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, args.steps,
                      args.warmup, False, graph=True)
    ms = r32["wall"] / args.steps * 1e3
    value = N ** 3 * args.steps / r32["wall"] / 1e6
    stats = r32["stats"]
    dom = max(stats, key=lambda k: stats[k]["seconds"])
    d = stats[dom]
    achieved = d["bytes"] / d["seconds"] / 1e9
    M, K = modal_sizes(alpha)
    decomp_note = ""
    if args.template == "scalari16":
        _, corr16, nz16 = scalari16_alpha(ctx, cfg["PrimordialIndex"], args.box, N)
        decomp_note = f": {nz16} nonzero coefficients, correlation {corr16:.6f}"
    out = {
        "metric": "Mcells/s for NG-2LPT IC at Nmesh=512, FP32 vs FP64; p50/p99 job latency",
        "value": round(value, 3), "unit": "Mcells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Gaussian field, lattice glass)",
        "config": {"workload": f"NG-2LPT IC, Nmesh=Npart={N}^3, Lbox={args.box:g} Mpc/h, "
                               f"{TEMPLATES[args.template]}{decomp_note} (M={M}, K={K}), EH P(k), fNL=100, "
                               f"seed 1234",
                   "precision": "fp32", "schedule": "batched"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(dom, N, args.template),
                     "algorithmic_bytes_per_launch": round(d["bytes"] / d["launches"]),
                     "avg_launch_us": round(d["seconds"] / d["launches"] * 1e6, 2)},
        "kernel_time_share": {k: round(v["seconds"] / r32["device"] * args.steps / args.steps, 4)
                              for k, v in stats.items()},
        "job_roofline": {"bytes_per_job": survey_bytes(N, M, K, 4),
                         "frac": round(survey_bytes(N, M, K, 4) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)},
    }
    if rg is not None:
        out["hipgraph"] = {"ms_per_step": round(rg["wall"] / args.steps * 1e3, 3),
                           "value": round(N ** 3 * args.steps / rg["wall"] / 1e6, 3)}
    if not args.no_fp64:
        r64 = time_job(aar, ctx, cfg, alpha, x0, glass, gbox, aar.AAR_FP64, aar.AAR_SCHED_PLAIN,
                       args.fp64_steps, 1, False, box=eb)
        v64 = N ** 3 * args.fp64_steps / r64["wall"] / 1e6
        out["fp64_plain"] = {"value": round(v64, 3), "unit": "Mcells/s",
                             "ms_per_step": round(r64["wall"] / args.fp64_steps * 1e3, 3),
                             "fp32_over_fp64": round(value / v64, 3)}
        # accuracy of the timed FP32 path: relative L2 error against the FP64 pairwise path on a
        # box^3 corner of every displacement field and of the non-Gaussian potential
        err = {n: float(np.linalg.norm(r32["box"][n] - r64["box"][n]) / np.linalg.norm(r64["box"][n]))
               for n in ERR_FIELDS}
        out["fp32_rel_l2_vs_fp64"] = {"sub_box": eb, "max": float(f"{max(err.values()):.3g}"),
                                      "fields": {n: float(f"{v:.3g}") for n, v in err.items()}}
    out["autotune"] = {"xy_tile_variant": {"x": tvar[0], "y": tvar[1]},
                       "us_per_launch_8_fields": {ax: [round(v * 1e6, 1) if v > 0 else None for v in tsec[i]]
                                                  for i, ax in enumerate(("x", "y"))},
                       "variants": "0: 128-B tile columns, 1: 64-B, 2: 256-B, 3: 128-B, four tiles per workgroup with the next tile's loads in flight (same radix plan)",
                       "xy_pair_group_cols": xy_group}
    out["copy_peak_gbs"] = round(copy_peak(ctx, min(N, 512)), 1)
    if not args.no_serve:
        out["serving"] = serving(aar, ctx, args.serve_jobs)
    if not args.no_decomp:
        out["decomposition"] = decomposition(aar, ctx)
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.template, ctx, alpha)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
