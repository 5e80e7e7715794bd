"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Tolerances: FP64 path 1e-10 relative (L2) on spectra and displacement fields;
particle positions and velocities are float32 and agree to 1 ulp (positions
that wrap at the box edge excepted); FP32 path <= 1e-5 relative L2.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def tables_for(cfg, alpha, x0, glass, gbox):
    tk = tt = None
    if cfg.get("WhichTransfer") == 2:
        tk, tt = cfg["transfer_k"], cfg["transfer_t"]
    return aar.tables(alpha, x0, glass, gbox, tk, tt)


def test_device_ranlxd_matches_oracle_bitwise(ctx):
    for seed in [1, 1234, 2 ** 31 - 1]:
        dev = ctx.ranlxd(seed, 1000)
        ref = orc.ranlxd_draws(seed, 1000)
        np.testing.assert_array_equal(dev, ref)


@pytest.mark.parametrize("prec,tol", [(aar.AAR_FP64, 1e-12), (aar.AAR_FP32, 2e-6)])
@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_fft_matches_numpy(ctx, N, prec, tol):
    rng = np.random.default_rng(N)
    H = N // 2 + 1
    spec = rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))
    ref = np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    out = ctx.fft3d(spec, +1, prec)
    assert rel_l2(out, ref) < tol
    real = rng.standard_normal((N, N, N))
    ref2 = np.fft.rfftn(real)
    out2 = ctx.fft3d(real, -1, prec)
    assert rel_l2(out2, ref2) < tol


@pytest.mark.parametrize("kw", [dict(), dict(SphereMode=1, Nsample=24), dict(PhaseFlip=1),
                                dict(FixedAmplitude=1, Nsample=20)])
def test_gaussian_field_vs_oracle(ctx, kw):
    cfg = cases.base_config(**{**dict(Nmesh=32, Nmax=32, Nsample=32), **kw})
    alpha, x0, glass, gbox = cases.job_inputs(cfg)
    t = tables_for(cfg, alpha, x0, glass, gbox)
    dev = ctx.gaussian_field(aar.params_from_dict(cfg), t, aar.AAR_FP64)
    ref = orc.gaussian(cfg)
    np.testing.assert_array_equal(dev == 0, ref == 0)
    assert np.max(np.abs(dev - ref)) <= 1e-13 * np.max(np.abs(ref))


def _compare_job(ctx, cfg, rows, x0, prec, sched, tol_fields):
    alpha, x0v, glass, gbox = cases.job_inputs(cfg, rows, x0)
    ref = orc.run(cfg, alpha, x0v, glass, gbox)
    t = tables_for(cfg, alpha, x0v, glass, gbox)
    job = aar.Job(ctx, aar.params_from_dict(cfg), t, prec, sched)
    job.run()
    job.sync()
    pos, vel = job.fetch()
    res = dict(phig=rel_l2(job.field("phig_spec"), ref["phig_spec"]),
               phi=rel_l2(job.field("phi_spec"), ref["phi_spec"]),
               phing_real=rel_l2(job.field("phing_real"), ref["phing_real"]))
    N = cfg["Nmesh"]
    # displacement fields: reconstruct the oracle's from its particle-independent pipeline
    for name in ["phig", "phi", "phing_real"]:
        assert res[name] < tol_fields, (name, res)
    box = cfg["Box"]
    names = ["disp_x", "disp_y", "disp_z", "disp2_x", "disp2_y", "disp2_z"]
    ref_disp = orc.lpt_fields(cfg, ref["phi_spec"])
    for name, rd in zip(names, ref_disp):
        res[name] = rel_l2(job.field(name), rd)
    res["vel"] = rel_l2(vel, ref["vel"])
    q = orc.glass_tile(cfg, glass, gbox).astype(np.float64)
    d_dev = (pos - q + box / 2) % box - box / 2
    d_ref = (ref["pos"] - q + box / 2) % box - box / 2
    res["disp_particles"] = rel_l2(d_dev, d_ref)
    for name in names:
        assert res[name] < tol_fields, (name, res)
    if prec == aar.AAR_FP64:
        dpos = np.abs(pos.astype(np.float64) - ref["pos"])
        dpos = np.minimum(dpos, box - dpos)
        ulp = np.spacing(np.abs(ref["pos"]).astype(np.float32)).astype(np.float64)
        assert np.all(dpos <= ulp), float((dpos / ulp).max())
        vulp = np.spacing(np.abs(ref["vel"]).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(vel - ref["vel"]) <= vulp), float((np.abs(vel - ref["vel"]) / vulp).max())
    else:
        assert res["vel"] < tol_fields, res
    return res


@pytest.mark.parametrize("sched", [aar.AAR_SCHED_PLAIN, aar.AAR_SCHED_BATCHED])
@pytest.mark.parametrize("table", ["local", "equilateral"])
def test_job_fp64_vs_oracle(ctx, sched, table):
    cfg = cases.base_config(Nmesh=32, Nsample=32, Nmax=32, N_modes=4)
    rows = cases.local_rows() if table == "local" else cases.equilateral_rows()
    _compare_job(ctx, cfg, rows, None, aar.AAR_FP64, sched, 1e-10)


@pytest.mark.parametrize("sched", [aar.AAR_SCHED_PLAIN, aar.AAR_SCHED_BATCHED])
def test_job_fp64_shipped_table_tabulated_transfer(ctx, sched):
    tk, tt = cases.load_transfer()
    cfg, rows, x0 = cases.shipped_table_config(N=32, WhichTransfer=2, transfer_k=tk, transfer_t=tt,
                                               Box=512.0)
    _compare_job(ctx, cfg, rows, x0, aar.AAR_FP64, sched, 1e-10)


def test_job_fp64_sphere_mode_cut(ctx):
    cfg = cases.base_config(Nmesh=32, Nsample=20, Nmax=32, N_modes=4, SphereMode=1, Seed=77)
    _compare_job(ctx, cfg, cases.equilateral_rows(), None, aar.AAR_FP64, aar.AAR_SCHED_BATCHED, 1e-10)


def test_job_c1_local_64(ctx):
    """Plain FP64 CPU-runnable case: 64^3, L=1000, EH, local fNL=100, 8 modes, seed 1234."""
    cfg = cases.base_config()
    _compare_job(ctx, cfg, cases.local_rows(), None, aar.AAR_FP64, aar.AAR_SCHED_PLAIN, 1e-10)


@pytest.mark.parametrize("table", ["equilateral", "shipped"])
def test_job_fp32_error_vs_oracle(ctx, table):
    if table == "shipped":
        cfg, rows, x0 = cases.shipped_table_config(N=64)
    else:
        cfg = cases.base_config(N_modes=4)
        rows, x0 = cases.equilateral_rows(), None
    res = _compare_job(ctx, cfg, rows, x0, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, 1e-5)
    print("fp32 rel L2", table, res)


@pytest.mark.parametrize("N,group", [(32, None), (64, 16)])
def test_graph_replays_match_eager(ctx, N, group):
    """The captured hipGraph gives the eager result on every replay, with eager work between;
    at 64^3 with the x / y pass pairs pinned to 16-column groups (two per field, field by field,
    and the fused z pass)."""
    lib = aar.lib()
    saved = lib.aar_xy_group_cols(N, aar.AAR_FP32)
    if group is not None:
        assert lib.aar_set_xy_group_cols(N, aar.AAR_FP32, group) == 0
    try:
        _graph_vs_eager(ctx, N)
    finally:
        lib.aar_set_xy_group_cols(N, aar.AAR_FP32, saved)


def _graph_vs_eager(ctx, N):
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=4)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.equilateral_rows())
    t = tables_for(cfg, alpha, x0, glass, gbox)
    eager = aar.Job(ctx, aar.params_from_dict(cfg), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    eager.run()
    eager.sync()
    pos0, vel0 = eager.fetch()
    phi0 = eager.field("phi_spec")
    g = aar.Job(ctx, aar.params_from_dict(cfg), t, aar.AAR_FP32, aar.AAR_SCHED_BATCHED)
    g.set_graph(True)
    for rep in range(3):
        g.run()
        g.sync()
        pos, vel = g.fetch()
        np.testing.assert_array_equal(pos, pos0)
        np.testing.assert_array_equal(vel, vel0)
        np.testing.assert_array_equal(g.field("phi_spec"), phi0)
        eager.run()  # unrelated eager work between replays
        eager.sync()
    g.close()
    eager.close()


AAR_ERR_BAD_ARG = 100001  # include/aarambam.h


def test_xy_pair_column_groups_give_identical_results(ctx):
    """The x / y pass pairs run over all fields at once or field by field in kz column groups
    (96 columns at 512^3 leaves a narrower last group; 256 is a whole field): bit-identical 3-D FFTs, and a batched
    FP64 job at 64^3 in two 16-column groups matches the same job with whole passes bitwise
    (and the oracle).  The setter rejects widths that are not multiples of 16 or exceed N/2."""
    lib = aar.lib()
    N = 512
    saved = lib.aar_xy_group_cols(N, aar.AAR_FP32)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((N, N, N))
    outs = []
    for gc in (0, 256, 128, 96):
        assert lib.aar_set_xy_group_cols(N, aar.AAR_FP32, gc) == 0
        assert lib.aar_xy_group_cols(N, aar.AAR_FP32) == gc
        y = ctx.fft3d(x, -1, aar.AAR_FP32)
        outs.append((y, ctx.fft3d(y, +1, aar.AAR_FP32)))
    assert lib.aar_set_xy_group_cols(N, aar.AAR_FP32, saved) == 0
    for y, z in outs[1:]:
        np.testing.assert_array_equal(y, outs[0][0])
        np.testing.assert_array_equal(z, outs[0][1])
    assert lib.aar_set_xy_group_cols(N, aar.AAR_FP32, 8) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_group_cols(N, aar.AAR_FP32, 512) == AAR_ERR_BAD_ARG
    assert lib.aar_set_xy_group_cols(100, aar.AAR_FP32, 16) == AAR_ERR_BAD_ARG

    cfg = cases.base_config(Nmesh=64, Nsample=64, Nmax=64, N_modes=12)
    alpha, x0, glass, gbox = cases.job_inputs(cfg, cases.equilateral_rows())
    t = tables_for(cfg, alpha, x0, glass, gbox)
    saved = lib.aar_xy_group_cols(64, aar.AAR_FP64)
    res = []
    for gc in (0, 16):
        assert lib.aar_set_xy_group_cols(64, aar.AAR_FP64, gc) == 0
        job = aar.Job(ctx, aar.params_from_dict(cfg), t, aar.AAR_FP64, aar.AAR_SCHED_BATCHED)
        job.run()
        job.sync()
        res.append((job.fetch(), job.field("phing_real"), job.field("disp2_x")))
        job.close()
    assert lib.aar_set_xy_group_cols(64, aar.AAR_FP64, saved) == 0
    for a, b in zip(res[0][0] + res[0][1:], res[1][0] + res[1][1:]):
        np.testing.assert_array_equal(a, b)
    ref = orc.run(cfg, alpha, x0, glass, gbox)
    assert rel_l2(res[1][1], ref["phing_real"]) < 1e-10


@pytest.mark.parametrize("N", [256, 512])
def test_xy_tile_shapes_give_identical_transforms(ctx, N):
    """Every x / y tile shape shares the radix plan: bit-identical FFTs, both directions.  At 256^3
    shape 3 (four tiles per workgroup) is not built and runs as shape 0; at 512^3 every shape
    runs."""
    lib = aar.lib()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((N, N, N))
    saved = [lib.aar_xy_tile_variant(N, aar.AAR_FP32, a) for a in (0, 1)]
    outs = []
    for v in range(4):
        for a in (0, 1):
            assert lib.aar_set_xy_tile_variant(N, aar.AAR_FP32, a, v) == 0
        y = ctx.fft3d(x, -1, aar.AAR_FP32)
        outs.append((y, ctx.fft3d(y, +1, aar.AAR_FP32)))
    for a in (0, 1):
        lib.aar_set_xy_tile_variant(N, aar.AAR_FP32, a, saved[a])
    for y, z in outs[1:]:
        np.testing.assert_array_equal(y, outs[0][0])
        np.testing.assert_array_equal(z, outs[0][1])
    sec, chosen = ctx.autotune(N, aar.AAR_FP32, nf=2)
    assert all(0 <= c < 4 for c in chosen)
    assert (sec[:, 0] > 0).all()
    assert (sec[:, 3] > 0).all() == (N >= 512)
