"""A prepared job re-run with a new seed and fNL (aar_job_reseed) equals a fresh job of that seed
and fNL, bit for bit, also when the new seed table is built while the previous run is still on
the device, and in graph mode; jobs with the same tables share one device-resident plan."""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def inputs(N=64, seed=5, fnl=100.0, rows=None):
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Seed=seed, Fnl=fnl)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows() if rows is None else rows)
    return cfg, aar.tables(alpha, x0, glass, gbox)


def fresh(ctx, cfg, t, prec, sched):
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.run()
    job.sync()
    out = job.fetch() + (job.field("phi_spec"),)
    job.close()
    return out


@pytest.mark.parametrize("prec,sched,graph", [(aar.AAR_FP32, aar.AAR_SCHED_BATCHED, False),
                                              (aar.AAR_FP32, aar.AAR_SCHED_BATCHED, True),
                                              (aar.AAR_FP64, aar.AAR_SCHED_PLAIN, False)])
def test_reseed_equals_fresh_job(ctx, prec, sched, graph):
    cfg, t = inputs()
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.set_graph(graph)
    job.run()  # seed 5, fNL 100: left running while the next seed table is built
    for seed, fnl in [(77, 100.0), (78, -42.5), (2 ** 31 - 1, -42.5)]:
        job.reseed(seed, fnl)
        job.run()
        job.sync()
        got = job.fetch() + (job.field("phi_spec"),)
        want = fresh(ctx, dict(cfg, Seed=seed, Fnl=fnl), t, prec, sched)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b, err_msg=f"seed {seed} fnl {fnl}")
    job.close()


def test_reseed_back_to_back_without_waiting(ctx):
    """Several reseeds and runs enqueued with no wait between them: the last result is the last
    seed's, and the seed tables of the two staging buffers never mix."""
    cfg, t = inputs(seed=1)
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    for seed in range(2, 9):
        job.reseed(seed, 10.0 * seed)
        job.run()
    job.sync()
    got = job.fetch()
    want = fresh(ctx, dict(cfg, Seed=8, Fnl=80.0), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    job.close()


def test_jobs_share_device_resident_tables():
    c = aar.Context(0)
    try:
        assert c.cached_plans() == (0, 0)
        cfg, t = inputs(seed=3)
        a = aar.Job(c, aar.params_from_dict(cfg), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        b = aar.Job(c, aar.params_from_dict(dict(cfg, Seed=4, Fnl=-7.0)), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        n, nbytes = c.cached_plans()
        assert n == 1 and nbytes > 12 * 64 ** 3  # the glass alone is 12 bytes per particle
        cfg2, t2 = inputs(seed=3, rows=cases.local_rows())  # another table: another plan
        d = aar.Job(c, aar.params_from_dict(cfg2), t2, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        e = aar.Job(c, aar.params_from_dict(cfg), t, aar.AAR_FP64, aar.AAR_SCHED_PLAIN)  # another schedule
        assert c.cached_plans()[0] == 3
        # a copy of the same tables in other host memory finds the cached plan
        cfg3, t3 = inputs(seed=9)
        f = aar.Job(c, aar.params_from_dict(cfg3), t3, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        assert c.cached_plans()[0] == 3
        for j in (a, b, d, e, f):
            j.run()
            j.sync()
        ref = fresh(c, cfg3, t3, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
        np.testing.assert_array_equal(f.fetch()[0], ref[0])
        for j in (a, b, d, e, f):
            j.close()
    finally:
        c.close()
