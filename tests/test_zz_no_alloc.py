"""aar_job_run_noalloc, the entry point that allocates nothing: through the C-ABI it gives the
particles of aar_job_prepare + aar_job_run bit for bit, on two calls in a row (the second with a
smaller mesh, so the shared scratch holds the first call's data), and refuses a job larger than
its scratch with AAR_ERR_SCRATCH.  The output arrays are device memory from torch."""
import ctypes as C

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")
torch = pytest.importorskip("torch")

AAR_ERR_SCRATCH = 100007  # include/aarambam.h


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def job_case(N, seed, rows, nm=12, **kw):
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=nm, Seed=seed, **kw)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, rows)
    return cfg, aar.tables(alpha, x0, glass, gbox)


def reference(ctx, cfg, t, prec, sched):
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.run()
    job.sync()
    out = job.fetch()
    job.close()
    return out


def noalloc(ctx, cfg, t, prec, sched):
    npart = cfg["GlassTileFac"] ** 3 * t.nglass
    pos = torch.full((npart, 3), float("nan"), dtype=torch.float32, device="cuda")
    vel = torch.full((npart, 3), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    p = aar.params_from_dict(cfg)
    rc = aar.lib().aar_job_run_noalloc(ctx.h, C.byref(p), C.byref(t.c), prec, sched, C.c_void_p(pos.data_ptr()),
                                       C.c_void_p(vel.data_ptr()), None)
    return rc, pos.cpu().numpy(), vel.cpu().numpy()


CASES = [  # (N, seed, table, precision, schedule): the second is smaller than the first
    (64, 21, "le", aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (32, 22, "equil", aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (64, 23, "le", aar.AAR_FP64, aar.AAR_SCHED_BATCHED),
    (32, 24, "local", aar.AAR_FP64, aar.AAR_SCHED_PLAIN),
]


def rows_of(name):
    return {"le": cases.le_rows, "equil": cases.equilateral_rows, "local": cases.local_rows}[name]()


def test_noalloc_matches_prepared_job_twice_in_a_row(ctx):
    for N, seed, table, prec, sched in CASES:
        cfg, t = job_case(N, seed, rows_of(table), Fnl=-60.0)
        want = reference(ctx, cfg, t, prec, sched)
        rc, pos, vel = noalloc(ctx, cfg, t, prec, sched)
        assert rc == 0, (N, table, rc)
        np.testing.assert_array_equal(pos, want[0], err_msg=f"{N} {table} {prec} {sched}")
        np.testing.assert_array_equal(vel, want[1], err_msg=f"{N} {table} {prec} {sched}")


def test_noalloc_refuses_jobs_beyond_its_scratch(ctx):
    cfg, t = job_case(128, 5, cases.local_rows(), GlassTileFac=2)  # a 64^3 glass tiled twice
    rc, pos, _ = noalloc(ctx, cfg, t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    assert rc == AAR_ERR_SCRATCH
    assert np.isnan(pos).all()  # nothing was written
    assert "scratch" in aar.lib().aar_status_string(rc).decode()
