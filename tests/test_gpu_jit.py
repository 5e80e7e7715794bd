"""The NG sum's fused z pass compiled at run time for a job's own product schedule (aar_jit.cpp)
against the library's kernel, which reads the same schedule from records: the same operations
in the same order, so particles and fields must be bit-identical, eagerly and under hipGraph
replay, for the 16-mode benchmark table (M = 16, K = 15), the shipped 15-mode table and the
equilateral table, in both precisions."""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")

FIELDS = ["phi_spec", "phing_real", "disp_x", "disp2_z"]


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def inputs(table, N):
    if table == "scalari16":
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=16, Box=2000.0 * N / 512,
                                GlassTileFac=max(1, N // 64), Seed=2024)
        rows, x0 = cases.scalari16_rows(), None
    elif table == "shipped":
        cfg, rows, x0 = cases.shipped_table_config(N, GlassTileFac=max(1, N // 64), Seed=2025)
    else:
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, GlassTileFac=max(1, N // 64), Seed=2026)
        rows, x0 = cases.equilateral_rows(), None
    alpha, x0v, glass, gbox = cases.job_inputs(cfg, rows, x0, glass_n=N // cfg["GlassTileFac"])
    return aar.params_from_dict(cfg), aar.tables(alpha, x0v, glass, gbox)


def run(ctx, p, t, prec, jit, graph=False):
    ctx.set_jit(jit)
    job = aar.Job(ctx, p, t, prec, aar.AAR_SCHED_BATCHED)
    used = job.jit()
    job.set_graph(graph)
    job.run()
    job.sync()
    if graph:  # a replay, with eager work of another job before it
        job.run()
        job.sync()
    out = {f: job.field(f) for f in FIELDS}
    out["pos"], out["vel"] = job.fetch()
    job.close()
    ctx.set_jit(True)
    return used, out


@pytest.mark.parametrize("table,N,prec", [("scalari16", 64, aar.AAR_FP32), ("scalari16", 64, aar.AAR_FP64),
                                          ("shipped", 128, aar.AAR_FP32), ("equilateral", 64, aar.AAR_FP32),
                                          ("scalari16", 256, aar.AAR_FP32)])
def test_compiled_schedule_equals_library_kernel(ctx, table, N, prec):
    print(f"jit {table} N={N} prec={prec}", flush=True)
    p, t = inputs(table, N)
    used_lib, ref = run(ctx, p, t, prec, False)
    used_jit, out = run(ctx, p, t, prec, True)
    assert used_lib == 0 and used_jit == 1
    for k in ref:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_compiled_schedule_under_graph_replay(ctx):
    print("jit graph", flush=True)
    p, t = inputs("scalari16", 64)
    _, ref = run(ctx, p, t, aar.AAR_FP32, False)
    used, out = run(ctx, p, t, aar.AAR_FP32, True, graph=True)
    assert used == 1
    for k in ref:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_other_schedules_keep_the_library_kernels(ctx):
    """Pairwise jobs and meshes that are not powers of two have no fused z pass to compile."""
    p, t = inputs("equilateral", 64)
    job = aar.Job(ctx, p, t, aar.AAR_FP64, aar.AAR_SCHED_PLAIN)
    assert job.jit() == 0
    job.close()
    p, t = inputs("equilateral", 96)
    job = aar.Job(ctx, p, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    assert job.jit() == 0
    job.close()
