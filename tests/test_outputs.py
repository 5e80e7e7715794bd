"""Output formats: the snapshot split over files as several writer tasks lay it
out (Aarambam/LPT/Basis/save.c:184-187, 252), the potential dump and its
collation (Aarambam/utils.py:307-347, with the dump's own Nmesh and StartX in
place of the file-name arithmetic)."""
import ctypes as C
import struct

import numpy as np
import pytest

import aarambam_amd as aar
from aarambam_amd import api
from tests import cases

FP = C.POINTER(C.c_float)


def read_gadget_file(path):
    b = open(path, "rb").read()
    off, blocks = 0, []
    while off < len(b):
        (m,) = struct.unpack_from("<i", b, off)
        body = b[off + 4: off + 4 + m]
        assert struct.unpack_from("<i", b, off + 4 + m)[0] == m
        blocks.append(body)
        off += 8 + m
    h = blocks[0]
    head = dict(npart=struct.unpack_from("6I", h, 0), mass=struct.unpack_from("6d", h, 24),
                ntot=struct.unpack_from("6I", h, 96), nfiles=struct.unpack_from("i", h, 124)[0])
    return (head, np.frombuffer(blocks[1], np.float32).reshape(-1, 3), np.frombuffer(blocks[2], np.float32).reshape(-1, 3),
            np.frombuffer(blocks[3], np.int32))


def _setup(tmp_path, n):
    d = dict(cases.base_config(Nmesh=16), OutputDir=str(tmp_path), FileBase="snap", NumFilesWrittenInParallel=1)
    p = api.params_from_dict(d)
    c = api.AarCosmo()
    t = api.tables(np.zeros(8 ** 3), np.zeros(16), cases.lattice_glass(4), 100.0)
    assert aar.lib().aar_cosmology(C.byref(p), C.byref(t.c), C.byref(c)) == 0
    rng = np.random.default_rng(1)
    pos = (rng.random((n, 3)) * 1000).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    return p, c, pos, vel


def _split(p, c, pos, vel, nfiles, max_record):
    got = C.c_int(-1)
    rc = aar.lib().aar_write_gadget_files(C.byref(p), C.byref(c), pos.ctypes.data_as(FP), vel.ctypes.data_as(FP),
                                          len(pos), None, nfiles, max_record, C.byref(got))
    return rc, got.value


@pytest.mark.parametrize("nfiles,max_record,expect", [(3, 0, 3), (0, 12 * 400, 3), (0, 12 * 333, 4), (7, 0, 7)])
def test_snapshot_split_over_files(tmp_path, nfiles, max_record, expect):
    n = 1000
    p, c, pos, vel = _setup(tmp_path, n)
    rc, got = _split(p, c, pos, vel, nfiles, max_record)
    assert rc == 0 and got == expect
    parts = [read_gadget_file(tmp_path / f"snap.{f}") for f in range(got)]
    assert not (tmp_path / "snap").exists() and not (tmp_path / f"snap.{got}").exists()
    for head, ps, _, _ in parts:
        assert head["nfiles"] == got and head["ntot"][1] == n and head["npart"][1] == len(ps)
        if max_record:
            assert 12 * len(ps) <= max_record
    np.testing.assert_array_equal(np.concatenate([q[1] for q in parts]), pos)
    np.testing.assert_array_equal(np.concatenate([q[2] for q in parts]), vel)
    np.testing.assert_array_equal(np.concatenate([q[3] for q in parts]), np.arange(1, n + 1))
    # the particle mass is that of the whole run in every file
    assert len({q[0]["mass"][1] for q in parts}) == 1


def test_one_file_split_equals_single_task_snapshot(tmp_path):
    p, c, pos, vel = _setup(tmp_path, 500)
    assert aar.lib().aar_write_gadget(C.byref(p), C.byref(c), pos.ctypes.data_as(FP), vel.ctypes.data_as(FP),
                                      500, None) == 0
    single = (tmp_path / "snap").read_bytes()
    (tmp_path / "snap").unlink()
    assert _split(p, c, pos, vel, 1, 0) == (0, 1)
    assert (tmp_path / "snap").read_bytes() == single


def test_split_refuses_impossible_counts(tmp_path):
    p, c, pos, vel = _setup(tmp_path, 500)
    assert _split(p, c, pos, vel, 2, 12 * 100)[0] != 0   # 250 per file > 100
    assert _split(p, c, pos, vel, 501, 0)[0] != 0        # more files than particles
    assert _split(p, c, pos, vel, 0, 11)[0] != 0         # not one particle per record


def _write_slab(path, N, start, nx, box, vals):
    W = 2 * (N // 2 + 1)
    rec = np.empty(nx * N * N, np.dtype([("ind", "<i4"), ("pot", "<f8")]))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(N), np.arange(N), indexing="ij")
    rec["ind"] = ((i * N + j) * W + k).reshape(-1)
    rec["pot"] = vals.reshape(-1)
    with open(path, "wb") as f:
        f.write(struct.pack("<iid", N * N * W, N, box))  # the reference stores task 0's size in every file
        f.write(rec.tobytes())


def test_collate_single_task_dumps(tmp_path):
    N = 8
    d = dict(cases.base_config(Nmesh=N), OutputDir=str(tmp_path), Seed=42)
    p = api.params_from_dict(d)
    rng = np.random.default_rng(3)
    g, ng = rng.standard_normal((2, N, N, N))
    dp = C.POINTER(C.c_double)
    assert aar.lib().aar_write_phi(C.byref(p), g.ctypes.data_as(dp), 0) == 0
    assert aar.lib().aar_write_phi(C.byref(p), ng.ctypes.data_as(dp), 1) == 0
    (tmp_path / "Gauss_potential_42.npy").write_bytes(b"not a dump")
    out = aar.collate_potential(str(tmp_path))
    np.testing.assert_array_equal(out["Gauss_potential"], g)
    np.testing.assert_array_equal(out["Nongauss_potential"], ng)
    assert sorted(f.name for f in tmp_path.iterdir()) == ["Gauss_potential_42.npy"]


def test_collate_slab_dumps(tmp_path):
    N = 12
    rng = np.random.default_rng(4)
    g = rng.standard_normal((N, N, N))
    # uneven slabs, written out of order; lexical order of the names is not x order past 9999
    for start, nx in [(7, 5), (0, 3), (3, 4)]:
        _write_slab(tmp_path / f"Gauss_potential_5_StartX_{start:04d}_NumX_{nx}", N, start, nx, 100.0,
                    g[start:start + nx])
    out = aar.collate_potential(str(tmp_path), remove=False)
    np.testing.assert_array_equal(out["Gauss_potential"], g)
    assert "Nongauss_potential" not in out
    _write_slab(tmp_path / "Gauss_potential_6_StartX_0000_NumX_12", N, 0, N, 100.0, g)
    with pytest.raises(ValueError, match="several seeds"):
        aar.collate_potential(str(tmp_path))
