"""Mesh sizes that are not powers of two on the GPU: the mixed-radix x / y and z passes
against numpy, and whole IC jobs at 48^3, 80^3, 96^3 and 112^3 (radices 3, 5 and 7) against
the oracle, FP64 pairwise and batched to 1e-10 (particles within one float ulp), FP32 batched
to 1e-5.  At these sizes the batched schedule runs its unfused form (z C2R, product kernel,
z R2C) and the pairwise 2LPT source."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import cases

pytestmark = pytest.mark.gpu

aar = pytest.importorskip("aarambam_amd")

from tests.test_gpu_parity import _compare_job  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = aar.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec,tol", [(aar.AAR_FP64, 1e-12), (aar.AAR_FP32, 2e-6)])
@pytest.mark.parametrize("N", [48, 80, 96, 112, 144, 240, 384])
def test_mixed_radix_fft_matches_numpy(ctx, N, prec, tol):
    print(f"fft N={N} prec={prec}", flush=True)
    rng = np.random.default_rng(N)
    H = N // 2 + 1
    spec = rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))
    ref = np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    assert rel_l2(ctx.fft3d(spec, +1, prec), ref) < tol
    del spec, ref
    real = rng.standard_normal((N, N, N))
    ref2 = np.fft.rfftn(real)
    assert rel_l2(ctx.fft3d(real, -1, prec), ref2) < tol


@pytest.mark.parametrize("N,table,sched", [
    (48, "local", aar.AAR_SCHED_PLAIN), (48, "equilateral", aar.AAR_SCHED_BATCHED),
    (80, "equilateral", aar.AAR_SCHED_PLAIN), (96, "equilateral", aar.AAR_SCHED_BATCHED),
    (112, "local", aar.AAR_SCHED_BATCHED)])
def test_job_fp64_vs_oracle(ctx, N, table, sched):
    print(f"job N={N} table={table} sched={sched}", flush=True)
    cfg = cases.base_config(Nmesh=N, Nsample=N, Nmax=N, N_modes=4, GlassTileFac=1, Seed=4321 + N)
    rows = cases.local_rows() if table == "local" else cases.equilateral_rows()
    _compare_job(ctx, cfg, rows, None, aar.AAR_FP64, sched, 1e-10)


def test_job_fp64_shipped_table_cube_cut_96(ctx):
    """The shipped 15-mode table with the reference example's cube cut (Nsample = Nmesh / 2)."""
    print("job N=96 shipped table", flush=True)
    cfg = cases.base_config(Nmesh=96, Nsample=48, Nmax=256, N_modes=15, PrimordialIndex=0.9660499, GlassTileFac=2)
    rows, x0 = cases.alpha_rows_file(), cases.x0_file(15)
    _compare_job(ctx, cfg, rows, x0, aar.AAR_FP64, aar.AAR_SCHED_BATCHED, 1e-10)


def test_job_fp32_vs_oracle_96(ctx):
    print("job N=96 fp32", flush=True)
    cfg = cases.base_config(Nmesh=96, Nsample=96, Nmax=96, N_modes=4, GlassTileFac=1)
    _compare_job(ctx, cfg, cases.equilateral_rows(), None, aar.AAR_FP32, aar.AAR_SCHED_BATCHED, 1e-5)
