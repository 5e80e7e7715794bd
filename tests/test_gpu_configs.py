"""The parity configurations at their full sizes.

* 256^3, Lbox = 1000 Mpc/h, local + equilateral table (12 modes), fNL = 100,
  seeds 1-4: the FP32 batched path against the FP64 pairwise path on the GPU
  (relative L2 <= 1e-5 on Phi, the six displacement fields and the
  velocities), and the FP64 path against the CPU oracle for seed 1 (1e-10).
* The synthetic dense 12-mode table (60 triples, 39 product pairs): FP64
  batched against the oracle at 64^3, FP32 against FP64 at 256^3.
* 1024^3, Lbox = 2000 Mpc/h, local + equilateral: FP32 against FP64 on the
  256^3 sub-box at the origin of the displacement and potential fields
  (the oracle cannot hold a 1024^3 grid; the FP64 GPU path is pinned to it
  by the smaller cases).
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


def run(ctx, cfg, rows, prec, sched, fields=("phi_spec",) + tuple(DISP)):
    alpha, x0, glass, gbox = cases.job_inputs(cfg, rows)
    job = aar.Job(ctx, aar.params_from_dict(cfg), aar.tables(alpha, x0, glass, gbox), prec, sched)
    job.run()
    job.sync()
    out = {f: job.field(f) for f in fields}
    out["pos"], out["vel"] = job.fetch()
    job.close()
    return out


def c2_config(seed, N=256, nm=12):
    return cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=nm, Box=1000.0, GlassTileFac=N // 64,
                             Seed=seed, Fnl=100.0)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_c2_le_256_fp32_vs_fp64(ctx, seed):
    cfg = c2_config(seed)
    r64 = run(ctx, cfg, cases.le_rows(), aar.AAR_FP64, aar.AAR_SCHED_PLAIN)
    r32 = run(ctx, cfg, cases.le_rows(), aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    err = {k: rel_l2(r32[k], r64[k]) for k in ["phi_spec", "vel"] + DISP}
    print("256^3 L+E seed", seed, "FP32 vs FP64 rel L2", err)
    assert max(err.values()) <= 1e-5, err


def test_c2_le_256_fp64_vs_oracle(ctx):
    cfg = c2_config(1)
    dev = run(ctx, cfg, cases.le_rows(), aar.AAR_FP64, aar.AAR_SCHED_PLAIN, fields=("phi_spec",))
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows())
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    assert rel_l2(dev["phi_spec"], ref["phi_spec"]) <= 1e-10
    assert rel_l2(dev["vel"], ref["vel"]) <= 1e-6  # float32 velocities


def test_dense12_fp64_batched_vs_oracle_64(ctx):
    cfg = c2_config(7, N=64)
    dev = run(ctx, cfg, cases.dense12_rows(), aar.AAR_FP64, aar.AAR_SCHED_BATCHED)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.dense12_rows())
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    assert rel_l2(dev["phi_spec"], ref["phi_spec"]) <= 1e-10
    for name, rd in zip(DISP, orc.lpt_fields(cfg, ref["phi_spec"])):
        assert rel_l2(dev[name], rd) <= 1e-10, name


def wide_rows(nm=24, n_triples=50, seed=24):
    """A table over 24 modes (more than the 16 mode fields the fused z pass holds in
    registers), so the batched schedule takes its unfused C2R / products / R2C passes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    triples = [(a, b, c) for a in range(nm) for b in range(a + 1) for c in range(b + 1)]
    pick = rng.choice(len(triples), size=n_triples, replace=False)
    return np.array([[*triples[i], rng.standard_normal()] for i in sorted(pick)])


@pytest.mark.parametrize("prec,tol", [(aar.AAR_FP64, 1e-10), (aar.AAR_FP32, 1e-5)])
def test_wide_table_beyond_the_fused_z_pass_vs_oracle_64(ctx, prec, tol):
    rows = wide_rows()
    cfg = c2_config(5, N=64, nm=24)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, rows)
    assert len({int(r[1]) for r in rows} | {int(r[2]) for r in rows}) > 16
    dev = run(ctx, cfg, rows, prec, aar.AAR_SCHED_BATCHED)
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    assert rel_l2(dev["phi_spec"], ref["phi_spec"]) <= tol
    for name, rd in zip(DISP, orc.lpt_fields(cfg, ref["phi_spec"])):
        assert rel_l2(dev[name], rd) <= tol, name


def test_dense12_256_fp32_vs_fp64(ctx):
    cfg = c2_config(8)
    r64 = run(ctx, cfg, cases.dense12_rows(), aar.AAR_FP64, aar.AAR_SCHED_PLAIN)
    r32 = run(ctx, cfg, cases.dense12_rows(), aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    err = {k: rel_l2(r32[k], r64[k]) for k in ["phi_spec", "vel"] + DISP}
    print("256^3 dense12 FP32 vs FP64 rel L2", err)
    assert max(err.values()) <= 1e-5, err


def test_field_box_matches_full_field(ctx):
    cfg = c2_config(3, N=64)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows())
    job = aar.Job(ctx, aar.params_from_dict(cfg), aar.tables(alpha, x0, glass, gbox), aar.AAR_FP32,
                  aar.AAR_SCHED_BATCHED)
    job.run()
    job.sync()
    for name in ["disp_y", "phing_real"]:
        full = job.field(name)
        np.testing.assert_array_equal(job.field_box(name, (5, 9, 3), (20, 17, 30)), full[5:25, 9:26, 3:33])
    job.close()


@pytest.mark.slow
def test_c4_le_1024_fp32_vs_fp64_on_256_subbox(ctx):
    N = 1024
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Box=2000.0, GlassTileFac=N // 64,
                            Seed=1234, Fnl=100.0)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows(), glass_n=64)
    names = DISP + ["phing_real"]
    boxes = {}
    for prec, sched in [(aar.AAR_FP64, aar.AAR_SCHED_PLAIN), (aar.AAR_FP32, aar.AAR_SCHED_BATCHED)]:
        job = aar.Job(ctx, aar.params_from_dict(cfg), aar.tables(alpha, x0, glass, gbox), prec, sched)
        job.run()
        job.sync()
        boxes[prec] = {n: job.field_box(n, (0, 0, 0), (256, 256, 256)) for n in names}
        job.close()
    err = {n: rel_l2(boxes[aar.AAR_FP32][n], boxes[aar.AAR_FP64][n]) for n in names}
    print("1024^3 L+E FP32 vs FP64 on the 256^3 sub-box, rel L2", err)
    assert max(err.values()) <= 1e-5, err
