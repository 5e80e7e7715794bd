"""Host side of the product library (no GPU): C-ABI exports, parameter file,
tables, cosmology, seed table and output writers, against the oracle and the
reference's file formats."""
import ctypes as C
import hashlib
import os
import re
import struct

import numpy as np
import pytest

import aarambam_amd as aar
from aarambam_amd import api
from oracle import pyoracle as orc
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "aarambam.h")).read()
    names = set(re.findall(r"\b(aar_[A-Za-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 25
    L = aar.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(api.EXPORTS) <= names


def test_host_ranlxd_known_answers():
    assert aar.lib().aar_ranlxd_nth(1, 397, 10000) == 3949287736
    assert aar.lib().aar_ranlxd_nth(1, 202, 10000) == 1998227290


@pytest.mark.parametrize("N,seed", [(16, 1), (64, 1234), (128, 2 ** 31 - 1)])
def test_seedtable_matches_oracle(N, seed):
    np.testing.assert_array_equal(api.seedtable(N, seed), orc.seedtable(N, seed))


@pytest.mark.parametrize("N,seed", [(96, 5), (512, 2 ** 31 - 1), (1024, 0), (1024, 77)])
def test_threaded_seedtable_matches_the_serial_stream(N, seed):
    """Threads that jump the ranlxd stream ahead to their first draw (shell i starts at draw 4 i^2:
    the jumps land on every register position the fill order reaches, after up to 87,000
    updates) give the oracle's serial table, whatever the thread count."""
    ref = orc.seedtable(N, seed)
    np.testing.assert_array_equal(api.seedtable(N, seed), ref)
    for t in (1, 2, 3, 8, 16):
        np.testing.assert_array_equal(api.seedtable(N, seed, threads=t), ref, err_msg=f"threads={t}")


def test_seedtable_threads_rejects_bad_arguments():
    out = np.empty(16, np.uint32)
    ptr = out.ctypes.data_as(C.POINTER(C.c_uint32))
    assert aar.lib().aar_seedtable_threads(4, 1, ptr, 0) == 100001
    assert aar.lib().aar_seedtable_threads(1, 1, ptr, 2) == 100001


@pytest.mark.parametrize("which", [1, 2])
def test_cosmology_bitwise_equal_to_oracle(which):
    tk, tt = cases.load_transfer()
    cfg = cases.base_config(WhichTransfer=which, transfer_k=tk, transfer_t=tt)
    ref = orc.cosmology(cfg)
    t = api.tables(np.zeros(27), np.zeros(6), cases.lattice_glass(4), 100.0, tk, tt)
    ours = api.cosmology(api.params_from_dict(cfg), t)
    for k, v in ref.items():
        assert ours[k] == v, (k, ours[k], v)


def test_read_params_reference_format(tmp_path):
    cfg = cases.base_config(WhichTransfer=2)
    pf = tmp_path / "LPTConfig"
    cases.write_param_file(str(pf), cfg, GlassFile="g.dat", TransferPath="t.dat", AlphaPath="a.dat",
                           AiPath="x.dat")
    p = aar.read_params(str(pf))
    d = api.params_to_dict(p)
    for k in ["Nmesh", "Nsample", "Nmax", "N_modes", "Seed", "WhichTransfer", "GlassTileFac"]:
        assert d[k] == cfg[k]
    for k in ["Box", "Omega", "OmegaBaryon", "Fnl", "PrimordialIndex", "Redshift"]:
        assert d[k] == pytest.approx(cfg[k], rel=1e-6)
    assert d["FileWithAlphaCoeff"] == "a.dat" and d["OutputDir"] == str(tmp_path)


def test_read_params_errors(tmp_path):
    cfg = cases.base_config()
    pf = tmp_path / "p"
    cases.write_param_file(str(pf), cfg, GlassFile="g", TransferPath="t", AlphaPath="a", AiPath="x")
    text = pf.read_text()
    (tmp_path / "missing").write_text(text.replace("Sigma8", "%Sigma8"))
    with pytest.raises(api.AarError, match="I miss a value for tag 'Sigma8'"):
        aar.read_params(str(tmp_path / "missing"))
    (tmp_path / "dup").write_text(text + "\nSeed 5\n")
    with pytest.raises(api.AarError, match="Tag 'Seed' not allowed or multiple defined"):
        aar.read_params(str(tmp_path / "dup"))
    with pytest.raises(api.AarError, match="not found"):
        aar.read_params(str(tmp_path / "nope"))


def test_load_tables_match_reference_semantics():
    Nm = 15
    a = np.empty(Nm ** 3)
    path = os.path.join(cases.GOLDEN, "AlphaTable.dat").encode()
    assert aar.lib().aar_load_alpha(path, Nm, a.ctypes.data_as(C.POINTER(C.c_double))) == 0
    ref = orc.alpha_table(Nm, cases.alpha_rows_file())
    np.testing.assert_array_equal(a.reshape(Nm, Nm, Nm), ref)
    x0 = np.empty(2 * Nm)
    assert aar.lib().aar_load_x0(os.path.join(cases.GOLDEN, "AiTable.dat").encode(), Nm,
                                 x0.ctypes.data_as(C.POINTER(C.c_double))) == 0
    np.testing.assert_array_equal(x0.reshape(Nm, 2), cases.x0_file(Nm))
    n = C.c_int(0)
    tpath = os.path.join(cases.GOLDEN, "TransferFunctionConverted.dat").encode()
    assert aar.lib().aar_load_transfer(tpath, C.byref(n), None, None) == 0
    k, t = np.empty(n.value), np.empty(n.value)
    assert aar.lib().aar_load_transfer(tpath, C.byref(n), k.ctypes.data_as(C.POINTER(C.c_double)),
                                       t.ctypes.data_as(C.POINTER(C.c_double))) == 0
    tk, tt = cases.load_transfer()
    np.testing.assert_array_equal(k, tk)
    np.testing.assert_array_equal(t, tt)


def test_table_errors():
    a = np.empty(27)
    assert aar.lib().aar_load_alpha(b"/nonexistent", 3, a.ctypes.data_as(C.POINTER(C.c_double))) == 17
    # a transfer table that does not reach k < 10^-4.6 (power.c:117-121)
    p = api.params_from_dict(cases.base_config(WhichTransfer=2))
    t = api.tables(np.zeros(27), np.zeros(6), cases.lattice_glass(4), 100.0, np.array([1e-3, 1.0]),
                   np.array([1.0, 0.5]))
    c = api.AarCosmo()
    assert aar.lib().aar_cosmology(C.byref(p), C.byref(t.c), C.byref(c)) == 1111
    # Nmax < Nsample is rejected before any allocation (main.c:200)
    p = api.params_from_dict(cases.base_config(Nmax=32, Nsample=64))
    t = api.tables(np.zeros(8 ** 3), np.zeros(16), cases.lattice_glass(4), 100.0)
    assert aar.lib().aar_job_bytes(C.byref(p), C.byref(t.c), 64, 1) == -1


def test_glass_file_format(tmp_path):
    path = tmp_path / "glass64.dat"
    pos = cases.write_lattice_glass(str(path), 64, 100.0)
    # byte-identical to the reference's shipped glass
    assert hashlib.sha256(path.read_bytes()).hexdigest() == cases.REFERENCE_GLASS_SHA256
    n, box = C.c_int(0), C.c_double(0)
    hdr = (C.c_uint8 * 256)()
    assert aar.lib().aar_read_glass(str(path).encode(), C.byref(n), None, C.byref(box), hdr) == 0
    assert n.value == 64 ** 3 and box.value == 100.0
    out = np.empty(3 * n.value, np.float32)
    assert aar.lib().aar_read_glass(str(path).encode(), C.byref(n), out.ctypes.data_as(C.POINTER(C.c_float)),
                                    C.byref(box), hdr) == 0
    np.testing.assert_array_equal(out.reshape(-1, 3), pos)
    assert aar.lib().aar_read_glass(str(tmp_path / "none").encode(), C.byref(n), None, None, None) == 121


def test_gadget_writer_layout(tmp_path):
    cfg = cases.base_config(Nmesh=16, Nsample=16, Nmax=16)
    d = dict(cfg, OutputDir=str(tmp_path), FileBase="snap", NumFilesWrittenInParallel=1)
    p = api.params_from_dict(d)
    t = api.tables(np.zeros(8 ** 3), np.zeros(16), cases.lattice_glass(4), 100.0)
    c = api.AarCosmo()
    assert aar.lib().aar_cosmology(C.byref(p), C.byref(t.c), C.byref(c)) == 0
    n = 1000
    rng = np.random.default_rng(0)
    pos = rng.random((n, 3), dtype=np.float32) * 1000
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    fp = C.POINTER(C.c_float)
    assert aar.lib().aar_write_gadget(C.byref(p), C.byref(c), pos.ctypes.data_as(fp),
                                      vel.ctypes.data_as(fp), n, None) == 0
    b = (tmp_path / "snap").read_bytes()
    off = 0

    def rec():
        nonlocal off
        (m,) = struct.unpack_from("i", b, off)
        body = b[off + 4: off + 4 + m]
        (m2,) = struct.unpack_from("i", b, off + 4 + m)
        assert m == m2
        off += 8 + m
        return body

    h = rec()
    assert len(h) == 256
    npart = struct.unpack_from("6I", h, 0)
    mass = struct.unpack_from("6d", h, 24)
    time, z = struct.unpack_from("2d", h, 72)
    ntot = struct.unpack_from("6I", h, 96)
    (nfiles,) = struct.unpack_from("i", h, 124)
    box, om, ol, hp = struct.unpack_from("4d", h, 128)
    assert npart[1] == n and ntot[1] == n and nfiles == 1 and box == cfg["Box"] and om == cfg["Omega"]
    expect_mass = cfg["Omega"] * 3 * c.Hubble * c.Hubble / (8 * np.pi * c.G) * cfg["Box"] ** 3 / n
    assert mass[1] == pytest.approx(expect_mass, rel=1e-14)
    assert time == pytest.approx(1 / (1 + cfg["Redshift"]), rel=1e-15)
    np.testing.assert_array_equal(np.frombuffer(rec(), np.float32).reshape(-1, 3), pos)
    np.testing.assert_array_equal(np.frombuffer(rec(), np.float32).reshape(-1, 3), vel)
    np.testing.assert_array_equal(np.frombuffer(rec(), np.int32), np.arange(1, n + 1))
    assert off == len(b)
    # more than one writer task is refused for a single task (save.c:112-117)
    p.NumFilesWrittenInParallel = 4
    assert aar.lib().aar_write_gadget(C.byref(p), C.byref(c), pos.ctypes.data_as(fp),
                                      vel.ctypes.data_as(fp), n, None) == 24131


def test_write_phi_layout(tmp_path):
    N = 8
    d = dict(cases.base_config(Nmesh=N), OutputDir=str(tmp_path), Seed=42)
    p = api.params_from_dict(d)
    f = np.arange(N ** 3, dtype=np.float64)
    assert aar.lib().aar_write_phi(C.byref(p), f.ctypes.data_as(C.POINTER(C.c_double)), 1) == 0
    b = (tmp_path / "Nongauss_potential_42").read_bytes()
    total, n = struct.unpack_from("2i", b, 0)
    (box,) = struct.unpack_from("d", b, 8)
    W = 2 * (N // 2 + 1)
    assert total == N * N * W + N * W and n == N and box == d["Box"]
    rec = np.frombuffer(b[16:], dtype=np.dtype([("c", "<i4"), ("v", "<f8")]))
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    np.testing.assert_array_equal(rec["c"], ((i * N + j) * W + k).reshape(-1))
    np.testing.assert_array_equal(rec["v"], f)


@pytest.mark.parametrize("which", [1, 2])
def test_cosmology_with_wdm_bitwise_equal_to_oracle(which):
    tk, tt = cases.load_transfer()
    cfg = cases.base_config(WhichTransfer=which, transfer_k=tk, transfer_t=tt, WDM_On=1, WDM_PartMass_in_kev=1.5)
    ref = orc.cosmology(cfg)
    t = api.tables(np.zeros(27), np.zeros(6), cases.lattice_glass(4), 100.0, tk, tt)
    ours = api.cosmology(api.params_from_dict(cfg), t)
    for k, v in ref.items():
        assert ours[k] == v, (k, ours[k], v)
    assert ref["Norm"] != orc.cosmology(dict(cfg, WDM_On=0))["Norm"]


def test_reciprocal_division_rounds_correctly(tmp_path):
    """The CIC readout divides by the glass box and the box as q = a rb with two FMA corrections
    (rb = RN(1/b), Markstein's iteration, aarambam_amd/csrc/aar_kernels.hip ddiv).  It must equal
    IEEE a / b bit for bit: 2e7 operands per divisor (uniform in [0, 4b), float-valued, and
    random exponents over 2^-30 .. 2^33) against the C library's division and fma."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    src = tmp_path / "ddiv.c"
    src.write_text(r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double ddiv(double a, double b, double rb) {
  double q = a * rb, r = fma(-b, q, a);
  q = fma(r, rb, q);
  r = fma(-b, q, a);
  return fma(r, rb, q);
}
int main(void) {
  const double bs[] = {100.0, 1000.0, 2000.0, 250.0, 3.0, 7.0, 1e-3, 123.456, 2000.0 / 3.0, 0.1};
  long long bad = 0;
  for (int ib = 0; ib < 10; ++ib) {
    const double b = bs[ib], rb = 1.0 / b;
    for (long long i = 0; i < 20000000LL; ++i) {
      uint64_t u = rnd();
      double a;
      if (i % 3 == 0) a = (double)(u >> 11) * 0x1p-53 * 4 * b;
      else if (i % 3 == 1) a = (double)(float)((double)(u >> 40) * 0x1p-24 * 4 * b);
      else { uint64_t bits = ((uint64_t)(1023 - 30 + (u >> 58)) << 52) | (rnd() >> 12); memcpy(&a, &bits, 8); }
      if (ddiv(a, b, rb) != a / b) ++bad;
    }
  }
  printf("%lld\n", bad);
  return 0;
}
""")
    exe = tmp_path / "ddiv"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=600)
    assert int(out.stdout.strip()) == 0


def test_job_refuses_glass_outside_its_box():
    """The CIC gather indexes the fields by each particle's unperturbed cell, so a glass
    coordinate outside [0, BoxSize], a non-finite one or a box that is not positive is refused
    before anything reaches the device (aar_job_bytes validates the job as aar_job_prepare does)."""
    cfg = cases.base_config(Nmesh=32, Nsample=32, Nmax=32, N_modes=4)
    alpha, x0, glass, gbox = cases.job_inputs(cfg)
    p = aar.params_from_dict(cfg)
    nbytes = lambda g, b: aar.lib().aar_job_bytes(p, api.tables(alpha, x0, g, b).c, aar.AAR_FP32,
                                                  aar.AAR_SCHED_BATCHED)
    assert nbytes(glass, gbox) > 0
    edge = glass.copy()
    edge[5] = gbox  # a coordinate on the upper face is the reference's clamped i == N case
    assert nbytes(edge, gbox) > 0
    for bad in (-1e-3, gbox * 1.01, np.nan, np.inf):
        g = glass.copy()
        g[7, 1] = bad
        assert nbytes(g, gbox) == -1, bad
    assert nbytes(glass, 0.0) == -1
    assert nbytes(glass, -100.0) == -1
    assert nbytes(glass, np.nan) == -1
