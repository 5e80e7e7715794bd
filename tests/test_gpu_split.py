"""GPU parity of the multi-output x / y pass pairs of the batched schedule.

The nine 2LPT spectra (main.c:1158-1277, 1311-1369) are one spectrum times factors of kx, ky and
kz, so the batched path writes only their five x-pass inputs (and the 2LPT displacements' two)
and the x pass stores each transform, times the ky and kz factors, to one, two or three fields.
Every column-group choice of these pairs must give bit-identical results, and the jobs must
match the oracle (FP64 1e-10 relative L2, particles within one float ulp; FP32 <= 1e-5).
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")

AAR_ERR_BAD_ARG = 100001  # include/aarambam.h
FIELDS = ["disp_x", "disp_y", "disp_z", "disp2_x", "disp2_y", "disp2_z", "phing_real"]


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def run_job(ctx, cfg, tabs, prec):
    job = aar.Job(ctx, aar.params_from_dict(cfg), tabs, prec, aar.AAR_SCHED_BATCHED)
    job.run()
    job.sync()
    out = {n: job.field(n) for n in FIELDS}
    out["pos"], out["vel"] = job.fetch()
    job.close()
    return out


@pytest.mark.parametrize("prec,N,groups", [(aar.AAR_FP64, 64, (0, 32, 16)),
                                           (aar.AAR_FP32, 128, (0, 64, 32, 16))])
def test_split_column_groups_give_identical_results(ctx, prec, N, groups):
    """Whole passes, all columns, halves, quarters: bitwise equal jobs, equal to the oracle."""
    lib = aar.lib()
    saved = lib.aar_xy_split_group_cols(N, prec)
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Seed=77)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows())
    tabs = aar.tables(alpha, x0, glass, gbox)
    res = []
    try:
        for gc in groups:
            assert lib.aar_set_xy_split_group_cols(N, prec, gc) == 0
            assert lib.aar_xy_split_group_cols(N, prec) == gc
            res.append(run_job(ctx, cfg, tabs, prec))
    finally:
        assert lib.aar_set_xy_split_group_cols(N, prec, saved) == 0
    for r in res[1:]:
        for n in res[0]:
            np.testing.assert_array_equal(r[n], res[0][n], err_msg=n)
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    ref_disp = orc.lpt_fields(cfg, ref["phi_spec"])
    tol = 1e-10 if prec == aar.AAR_FP64 else 1e-5
    for n, rd in zip(FIELDS[:6], ref_disp):
        assert rel_l2(res[0][n], rd) < tol, n
    assert rel_l2(res[0]["phing_real"], ref["phing_real"]) < tol
    box = cfg["Box"]
    if prec == aar.AAR_FP64:
        dpos = np.abs(res[0]["pos"].astype(np.float64) - ref["pos"])
        dpos = np.minimum(dpos, box - dpos)
        ulp = np.spacing(np.abs(ref["pos"]).astype(np.float32)).astype(np.float64)
        assert np.all(dpos <= ulp), float((dpos / ulp).max())
        vulp = np.spacing(np.abs(ref["vel"]).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(res[0]["vel"] - ref["vel"]) <= vulp)
    else:
        assert rel_l2(res[0]["vel"], ref["vel"]) < tol


@pytest.mark.parametrize("prec,N", [(aar.AAR_FP64, 64), (aar.AAR_FP32, 512)])
def test_split_y_tile_shapes_give_identical_results(ctx, prec, N):
    """The multi-output pairs' y pass has a tile shape of its own (axis 2 of the variant table):
    every shape, and -1 (follow the y pass), gives bitwise equal jobs.  At 512^3 FP32 every shape
    is built; the real fields are compared on an all-x, all-y slab of four z planes, which every
    (x, kz) column of the y pass reaches through the z transform."""
    lib = aar.lib()
    saved = lib.aar_xy_tile_variant(N, prec, 2)
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=12, Seed=91)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.le_rows())
    tabs = aar.tables(alpha, x0, glass, gbox)
    res = []
    try:
        for v in (-1, 0, 1, 2, 3):
            assert lib.aar_set_xy_tile_variant(N, prec, 2, v) == 0
            assert lib.aar_xy_tile_variant(N, prec, 2) == (lib.aar_xy_tile_variant(N, prec, 1) if v < 0 else v)
            job = aar.Job(ctx, aar.params_from_dict(cfg), tabs, prec, aar.AAR_SCHED_BATCHED)
            job.run()
            job.sync()
            if N <= 64:
                res.append({n: job.field(n) for n in FIELDS})
            else:
                res.append({n: job.field_box(n, (0, 0, 5), (N, N, 4)) for n in FIELDS[:6]})
            job.close()
    finally:
        lib.aar_set_xy_tile_variant(N, prec, 2, saved)
    assert lib.aar_set_xy_tile_variant(N, prec, 2, 4) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_tile_variant(N, prec, 2, -2) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_tile_variant(N, prec, 1, -1) == AAR_ERR_BAD_ARG
    assert np.abs(res[0]["disp2_x"]).max() > 0
    for r in res[1:]:
        for n in res[0]:
            np.testing.assert_array_equal(r[n], res[0][n], err_msg=n)


def test_split_group_setter_rejects_bad_widths():
    lib = aar.lib()
    assert lib.aar_set_xy_split_group_cols(512, aar.AAR_FP32, 8) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_split_group_cols(512, aar.AAR_FP32, 512) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_split_group_cols(512, aar.AAR_FP32, -2) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_split_group_cols(96, aar.AAR_FP32, 16) == AAR_ERR_BAD_ARG


def test_split_autotune_reports_every_choice(ctx):
    """One tuner search at 256^3 times whole passes, all columns, halves, quarters and eighths of
    the multi-output pairs, then the tile shapes of their y pass, and keeps one of each."""
    ts, ps, ss = ctx.autotune_report(256, aar.AAR_FP32, nf=2)
    assert np.all(ss[:8] > 0) and ss[8] == -1, ss  # y shape 3 is not built at 256^3
    assert aar.lib().aar_xy_tile_variant(256, aar.AAR_FP32, 2) in (0, 1, 2)
    assert np.all(ts[:, 0] > 0) and np.all(ps[0] > 0)
    assert aar.lib().aar_xy_split_group_cols(256, aar.AAR_FP32) in (0, 128, 64, 32, 16)


@pytest.mark.parametrize("prec,N", [(aar.AAR_FP64, 64), (aar.AAR_FP32, 128), (aar.AAR_FP32, 256)])
def test_mode_fields_filtered_in_the_x_pass_match_the_k_space_filter(prec, N):
    """The mode fields filtered in the loads of their x pass and written in k-space
    first by k_modes_spec: bitwise equal jobs (the same operations in the same order), with the
    shipped 15-mode table and the reference example's cube cut (Nsample = N/2)."""
    cfg = cases.base_config(Nmesh=N, Nsample=N // 2, Nmax=N, N_modes=15, Seed=5)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.alpha_rows_file())
    tabs = aar.tables(alpha, x0, glass, gbox)
    res = []
    for mode in (2, 0):  # always, never
        c = aar.Context(0)
        c.set_modes_x(mode)
        res.append(run_job(c, cfg, tabs, prec))
        c.close()
    for n in res[0]:
        np.testing.assert_array_equal(res[0][n], res[1][n], err_msg=n)
    if prec == aar.AAR_FP64 and N == 64:
        ref = orc.run(cfg, alpha, x0, glass, gbox)
        assert rel_l2(res[0]["phing_real"], ref["phing_real"]) < 1e-10
