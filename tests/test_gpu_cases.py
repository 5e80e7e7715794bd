"""The benchmark's modal table, the reference's own example run and 512^3 transforms on the GPU.

* The 16-mode ScalarI table of the 512^3 benchmark (M = 16 mode fields, K = 15 outer indices:
  the fused z pass at its register ceiling) as whole jobs at 64^3 against the oracle: FP64
  pairwise and batched to 1e-10, FP32 batched to 1e-5.
* The reference's example run (examples/MakeICs.py: shipped 15-mode tables, tabulated transfer,
  cube cut at Nsample = Nmesh / 2 with Nmax = Nsample, 64^3 glass tiled 4 times so the particles
  are an eighth of the mesh), scaled to 64^3 against the oracle, and at its full size (512^3 mesh,
  256^3 particles) FP32 batched under hipGraph against FP64 pairwise.
* The benchmark job itself (512^3, L = 2000, the 16-mode table) FP32 batched under hipGraph with
  the autotuned tile shapes against FP64 pairwise on a 128^3 sub-box.
* 512^3 R2C / C2R against numpy.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")

DISP = ["disp_x", "disp_y", "disp_z", "disp2_x", "disp2_y", "disp2_z"]


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def tables(cfg, alpha, x0, glass, gbox):
    tk = tt = None
    if cfg.get("WhichTransfer") == 2:
        tk, tt = cfg["transfer_k"], cfg["transfer_t"]
    return aar.tables(alpha, x0, glass, gbox, tk, tt)


def run_job(ctx, cfg, t, prec, sched, fields=("phi_spec",) + tuple(DISP), graph=False, box=0):
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.set_graph(graph)
    job.run()
    job.sync()
    if graph:  # a replay, not only the capture run
        job.run()
        job.sync()
    out = {}
    if box:
        out.update({n: job.field_box(n, (0, 0, 0), (box, box, box)) for n in DISP + ["phing_real"]})
    else:
        out.update({f: job.field(f) for f in fields})
    out["pos"], out["vel"] = job.fetch()
    job.close()
    return out


def check_vs_oracle(ctx, cfg, alpha, x0, glass, gbox, prec, sched, tol):
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    t = tables(cfg, alpha, x0, glass, gbox)
    dev = run_job(ctx, cfg, t, prec, sched)
    err = {"phi_spec": rel_l2(dev["phi_spec"], ref["phi_spec"]), "vel": rel_l2(dev["vel"], ref["vel"])}
    for name, rd in zip(DISP, orc.lpt_fields(cfg, ref["phi_spec"])):
        err[name] = rel_l2(dev[name], rd)
    vtol = tol if prec == aar.AAR_FP32 else 1e-6  # velocities are float32
    assert max(v for k, v in err.items() if k != "vel") <= tol, err
    assert err["vel"] <= vtol, err
    if prec == aar.AAR_FP64:  # positions within one float ulp, wrap-around excepted
        dpos = np.abs(dev["pos"].astype(np.float64) - ref["pos"])
        dpos = np.minimum(dpos, cfg["Box"] - dpos)
        assert np.all(dpos <= np.spacing(np.abs(ref["pos"]).astype(np.float32)))
    return err


def scalari16_inputs(N=64):
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=16, Box=2000.0 * N / 512, Seed=1234)
    alpha = orc.alpha_table(16, cases.scalari16_rows())
    x0 = cases.x0_closed_form(16, cfg["PrimordialIndex"], cfg["Box"], N)
    return cfg, alpha, x0, cases.lattice_glass(N, 100.0), 100.0


def test_scalari16_table_has_sixteen_mode_fields():
    rows = cases.scalari16_rows()
    modes = {int(r[1]) for r in rows} | {int(r[2]) for r in rows}
    outer = {int(r[0]) for r in rows}
    assert len(rows) == 132 and len(modes) == 16 and len(outer) == 15


@pytest.mark.parametrize("prec,sched,tol", [(aar.AAR_FP64, aar.AAR_SCHED_PLAIN, 1e-10),
                                            (aar.AAR_FP64, aar.AAR_SCHED_BATCHED, 1e-10),
                                            (aar.AAR_FP32, aar.AAR_SCHED_BATCHED, 1e-5)])
def test_scalari16_m16_job_vs_oracle_64(ctx, prec, sched, tol):
    err = check_vs_oracle(ctx, *scalari16_inputs(64), prec, sched, tol)
    print("scalari16 64^3", prec, sched, err)


@pytest.mark.parametrize("prec,sched,tol", [(aar.AAR_FP64, aar.AAR_SCHED_PLAIN, 1e-10),
                                            (aar.AAR_FP64, aar.AAR_SCHED_BATCHED, 1e-10),
                                            (aar.AAR_FP32, aar.AAR_SCHED_BATCHED, 1e-5)])
def test_reference_example_cut_vs_oracle_64(ctx, prec, sched, tol):
    cfg = cases.example_config(64)
    assert cfg["Nsample"] == 32 and cfg["Nmax"] == 32 and cfg["SphereMode"] == 0
    alpha = orc.alpha_table(15, cases.alpha_rows_file())
    x0 = cases.x0_closed_form(15, cfg["PrimordialIndex"], cfg["Box"], cfg["Nmax"])
    glass = cases.lattice_glass(8, 100.0)  # 8^3 glass tiled 4 times: 32^3 particles on a 64^3 mesh
    err = check_vs_oracle(ctx, cfg, alpha, x0, glass, 100.0, prec, sched, tol)
    print("example cut 64^3", prec, sched, err)


def fp32_graph_vs_fp64(ctx, cfg, alpha, x0, glass, gbox, box):
    t = tables(cfg, alpha, x0, glass, gbox)
    r64 = run_job(ctx, cfg, t, aar.AAR_FP64, aar.AAR_SCHED_PLAIN, box=box)
    r32 = run_job(ctx, cfg, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, graph=True, box=box)
    err = {n: rel_l2(r32[n], r64[n]) for n in DISP + ["phing_real", "vel"]}
    assert max(err.values()) <= 1e-5, err
    return err


def test_reference_example_512_fp32_graph_vs_fp64(ctx):
    cfg = cases.example_config(512)
    alpha = orc.alpha_table(15, cases.alpha_rows_file())
    x0 = cases.x0_file(15)  # the shipped AiTable is this config's (Nmax 256, n_s 0.9660499)
    err = fp32_graph_vs_fp64(ctx, cfg, alpha, x0, cases.lattice_glass(64, 100.0), 100.0, 128)
    print("example 512^3 (256^3 particles) FP32 graph vs FP64, 128^3 sub-box", err)


def test_benchmark_job_512_fp32_graph_autotuned_vs_fp64(ctx):
    cfg, alpha, x0, _, gbox = scalari16_inputs(512)
    cfg["GlassTileFac"] = 8
    ctx.autotune(512, aar.AAR_FP32, nf=8)
    err = fp32_graph_vs_fp64(ctx, cfg, alpha, x0, cases.lattice_glass(64, 100.0), gbox, 128)
    print("scalari16 512^3 FP32 graph vs FP64, 128^3 sub-box", err)


@pytest.mark.parametrize("prec,tol", [(aar.AAR_FP64, 1e-12), (aar.AAR_FP32, 2e-6)])
def test_fft_512_matches_numpy(ctx, prec, tol):
    N = 512
    rng = np.random.default_rng(512)
    real = rng.standard_normal((N, N, N))
    ref = np.fft.rfftn(real)
    assert rel_l2(ctx.fft3d(real, -1, prec), ref) < tol
    del real
    spec = ref * np.exp(2j * np.pi * rng.random(ref.shape))  # non-Hermitian on the kz = 0, N/2 planes
    ref2 = np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    assert rel_l2(ctx.fft3d(spec, +1, prec), ref2) < tol


@pytest.mark.parametrize("prec,tol", [(aar.AAR_FP64, 1e-12), (aar.AAR_FP32, 2e-6)])
def test_fft_512_radix_plans_match_numpy(ctx, prec, tol):
    """Each radix plan of the x / y passes (8 8 8, 16 16 2, 2 16 16 at 512) against numpy."""
    N = 512
    lib = aar.lib()
    saved = lib.aar_xy_radix_plan(N, prec)
    rng = np.random.default_rng(77)
    real = rng.standard_normal((N, N, N))
    ref = np.fft.rfftn(real)
    outs = []
    try:
        for plan in range(3):
            assert lib.aar_set_xy_radix_plan(N, prec, plan) == 0 and lib.aar_xy_radix_plan(N, prec) == plan
            out = ctx.fft3d(real, -1, prec)
            assert rel_l2(out, ref) < tol, plan
            outs.append(out)
    finally:
        lib.aar_set_xy_radix_plan(N, prec, saved)
    assert any(np.any(o != outs[0]) for o in outs[1:])  # the plans round differently
    assert lib.aar_set_xy_radix_plan(256, prec, 1) == 100001  # plans other than 0 exist at 512, 1024


def test_1024_job_radix_plans_agree(ctx):
    """The 1024^3 local + equilateral FP32 job with each radix plan: displacement and potential
    sub-boxes agree with plan 0's to the FP32 rounding of the transforms."""
    N = 1024
    lib = aar.lib()
    saved = lib.aar_xy_radix_plan(N, aar.AAR_FP32)
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Box=2000.0, GlassTileFac=16, Seed=99)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows(), glass_n=64)
    t = aar.tables(alpha, x0, glass, gbox)
    boxes = []
    try:
        for plan in range(3):
            assert lib.aar_set_xy_radix_plan(N, aar.AAR_FP32, plan) == 0
            boxes.append(run_job(ctx, cfg, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, box=64))
    finally:
        lib.aar_set_xy_radix_plan(N, aar.AAR_FP32, saved)
    for b in boxes[1:]:
        for n in DISP + ["phing_real"]:
            assert rel_l2(b[n], boxes[0][n]) < 2e-6, n
    sec, plan = ctx.autotune_plans(N, aar.AAR_FP32, nf=4)
    assert (sec > 0).all() and plan == int(np.argmin(sec))
