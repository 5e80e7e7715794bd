"""Serving loop (aar_server_*): continuous batching of IC jobs on one device.

Each job's particles must be bit-identical whether it runs alone (plain
aar_job_prepare/run, or a server with one slot) or overlapped with other jobs
of mixed Nmesh, table, seed, precision and schedule: the kernels are
deterministic (no atomics, fixed reduction order), so the stated tolerance is
exact equality.  Admission by HBM footprint is checked with budgets that fit
one job at a time.
"""
import time

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


def make_job(N, template, seed, fnl):
    if template == "scalari":
        cfg, rows, x0 = cases.shipped_table_config(N, Seed=seed, Fnl=fnl)
    elif template == "scalari16":
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=16, Box=2000.0 * N / 512, Seed=seed, Fnl=fnl,
                                GlassTileFac=max(1, N // 64))
        rows, x0 = cases.scalari16_rows(), None
    else:
        nm = 8 if template == "local" else 12
        cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=nm, Seed=seed, Fnl=fnl,
                                GlassTileFac=max(1, N // 64))
        rows, x0 = (cases.local_rows() if template == "local" else cases.equilateral_rows()), None
    # a lattice glass of at most 64^3 particles, tiled (GlassTileFac) to N^3
    alpha, x0, glass, gbox = cases.job_inputs(cfg, rows, x0, glass_n=N // cfg["GlassTileFac"])
    return aar.params_from_dict(cfg), aar.tables(alpha, x0, glass, gbox)


MIX = [  # (N, template, seed, fNL, precision, schedule)
    (32, "local", 11, 100.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (64, "equil", 12, -150.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (32, "scalari", 13, 37.5, aar.AAR_FP64, aar.AAR_SCHED_PLAIN),
    (64, "local", 14, 200.0, aar.AAR_FP64, aar.AAR_SCHED_BATCHED),
    (32, "equil", 15, -80.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (64, "scalari", 16, 120.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (128, "local", 17, -20.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),
    (96, "equil", 18, 64.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED),  # mixed-radix passes
    (48, "scalari", 19, -33.0, aar.AAR_FP64, aar.AAR_SCHED_PLAIN),
]


def alone(ctx, spec):
    N, tpl, seed, fnl, prec, sched = spec
    p, t = make_job(N, tpl, seed, fnl)
    job = aar.Job(ctx, p, t, prec, sched)
    job.run()
    job.sync()
    out = job.fetch()
    job.close()
    return out


def test_server_results_equal_alone_and_in_the_mix(ctx):
    ref = [alone(ctx, s) for s in MIX]
    jobs = [make_job(*s[:4]) for s in MIX]
    budget = sum(aar.lib().aar_job_bytes(p, t.c, s[4], s[5]) for (p, t), s in zip(jobs, MIX)) + (1 << 24)
    # one slot: the server path alone
    srv = aar.Server(ctx, budget, max_active=1)
    for (p, t), s, (rp, rv) in zip(jobs, MIX, ref):
        tk = srv.submit(p, t, s[4], s[5], keep=True)
        srv.wait(tk)
        pos, vel = srv.fetch(tk)
        srv.release(tk)
        np.testing.assert_array_equal(pos, rp)
        np.testing.assert_array_equal(vel, rv)
    srv.close()
    # four slots, everything submitted at once: the jobs overlap on the device
    srv = aar.Server(ctx, budget, max_active=4)
    tickets = [srv.submit(p, t, s[4], s[5], keep=True) for (p, t), s in zip(jobs, MIX)]
    for tk, s, (rp, rv) in zip(tickets, MIX, ref):
        lat, q = srv.wait(tk)
        assert lat >= q >= 0
        pos, vel = srv.fetch(tk)
        np.testing.assert_array_equal(pos, rp, err_msg=str(s))
        np.testing.assert_array_equal(vel, rv, err_msg=str(s))
        srv.release(tk)
    st = srv.stats()
    assert st["submitted"] == st["completed"] == len(MIX)
    assert st["peak_active"] >= 2
    assert st["peak_bytes"] <= budget
    srv.close()


def test_admission_by_hbm_footprint(ctx):
    specs = [(64, "local", 20 + q, 100.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED) for q in range(4)]
    jobs = [make_job(*s[:4]) for s in specs]
    need = aar.lib().aar_job_bytes(jobs[0][0], jobs[0][1].c, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    assert need > 0
    srv = aar.Server(ctx, need * 3 // 2, max_active=4)  # room for one job at a time
    tickets = [srv.submit(p, t, s[4], s[5]) for (p, t), s in zip(jobs, specs)]
    for tk in tickets:
        srv.wait(tk)
        srv.release(tk)
    st = srv.stats()
    assert st["completed"] == 4 and st["peak_active"] == 1 and st["peak_bytes"] <= need * 3 // 2
    # a job that can never fit is refused at submission
    big = make_job(128, "local", 1, 100.0)
    with pytest.raises(aar.AarError) as e:
        srv.submit(*big, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    assert e.value.code == 100003
    srv.close()


def test_kept_results_hold_their_memory_until_release(ctx):
    a = make_job(64, "equil", 31, 50.0)
    b = make_job(64, "local", 32, 50.0)
    need = aar.lib().aar_job_bytes(a[0], a[1].c, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    srv = aar.Server(ctx, need + need // 2, max_active=2)
    ta = srv.submit(*a, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, keep=True)
    srv.wait(ta)
    tb = srv.submit(*b, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, keep=True)
    time.sleep(0.3)
    assert srv.stats()["completed"] == 1  # b waits for a's piece of the arena
    ref = alone(ctx, (64, "equil", 31, 50.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED))
    pos, _ = srv.fetch(ta)
    np.testing.assert_array_equal(pos, ref[0])
    srv.release(ta)
    srv.wait(tb)
    ref_b = alone(ctx, (64, "local", 32, 50.0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED))
    np.testing.assert_array_equal(srv.fetch(tb)[0], ref_b[0])
    srv.release(tb)
    srv.close()


MIX_AT_SIZE = [  # (N, template, seed, fNL): the benchmark stream's sizes and templates, FP32 batched
    (512, "scalari16", 101, -120.0),
    (256, "equil", 102, 55.0),
    (128, "local", 103, 190.0),
    (512, "equil", 104, 80.0),
    (256, "scalari16", 105, -30.0),
    (128, "scalari16", 106, 12.5),
    (256, "local", 107, -200.0),
]


def test_server_at_benchmark_sizes_equal_alone_and_in_the_mix(ctx):
    """Jobs of 128^3, 256^3 and 512^3 with the local, equilateral and 16-mode ScalarI tables,
    all submitted at once to four slots: at least three run together, and each result is
    bit-identical to the same job run alone."""
    specs = [s + (aar.AAR_FP32, aar.AAR_SCHED_BATCHED) for s in MIX_AT_SIZE]
    jobs = [make_job(*s[:4]) for s in specs]
    budget = sum(aar.lib().aar_job_bytes(p, t.c, s[4], s[5]) for (p, t), s in zip(jobs, specs)) + (1 << 26)
    srv = aar.Server(ctx, budget, max_active=4)
    tickets = [srv.submit(p, t, s[4], s[5], keep=True) for (p, t), s in zip(jobs, specs)]
    for tk, s in zip(tickets, specs):
        srv.wait(tk)
        pos, vel = srv.fetch(tk)
        srv.release(tk)
        rp, rv = alone(ctx, s)
        np.testing.assert_array_equal(pos, rp, err_msg=str(s))
        np.testing.assert_array_equal(vel, rv, err_msg=str(s))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(specs)
    assert st["peak_active"] >= 3, st
