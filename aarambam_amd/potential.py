"""Collating the potential dumps into N^3 arrays.

Mirrors collate_potential of the reference (Aarambam/utils.py:307-347): for
Gauss_potential and Nongauss_potential it reads the dumps in OutputDir, returns
the grids as {name: float64 (N, N, N)} and removes the dump files.  The reference
takes Nmesh as the sum of the last '_' field of the file names, which is the
slab width NumX for several tasks' files (save.c:66-79) but the Seed for the one
file of a single task, and it orders slabs by the sorted names; here Nmesh is the
integer the dump's header stores after the record count (save.c:88-90), each
slab is placed by its StartX, and each record's coordinate is checked against
its position.  Dumps of different seeds in one directory raise instead of being
mixed.
"""
import glob
import os
import re

import numpy as np

_REC = np.dtype([("ind", "<i4"), ("pot", "<f8")])
_NAME = re.compile(r"^(Gauss|Nongauss)_potential_(-?\d+)(?:_StartX_(\d+)_NumX_(\d+))?$")


def _dump_files(out_dir, kind):
    files = []
    for f in glob.glob(os.path.join(out_dir, f"{kind}_potential_*")):
        m = _NAME.match(os.path.basename(f))
        if m and m.group(1) == kind:  # skips .npy and other files
            sx = int(m.group(3)) if m.group(3) is not None else None
            nx = int(m.group(4)) if m.group(4) is not None else None
            files.append((int(m.group(2)), sx, nx, f))
    return files


def collate_potential(OutputDir, remove=True):
    """{"Gauss_potential": phi_G, "Nongauss_potential": phi} as (N, N, N) float64 arrays."""
    out = {}
    for kind in ["Gauss", "Nongauss"]:
        files = _dump_files(OutputDir, kind)
        if not files:
            continue
        seeds = {s for s, *_ in files}
        if len(seeds) > 1:
            raise ValueError(f"{kind}_potential dumps of several seeds in {OutputDir}: {sorted(seeds)}")
        N = None
        slabs = []
        for _, sx, nx, f in files:
            with open(f, "rb") as fh:
                _total, n = np.frombuffer(fh.read(8), "<i4")
            if N is None:
                N = int(n)
            elif N != n:
                raise ValueError(f"{f}: Nmesh {n} differs from {N}")
            if sx is None:
                if len(files) > 1:
                    raise ValueError(f"single-task dump {f} next to slab dumps")
                sx, nx = 0, N
            slabs.append((sx, nx, f))
        slabs.sort()
        arr = np.empty((N, N, N), np.float64)
        W = 2 * (N // 2 + 1)
        x = 0
        for sx, nx, f in slabs:
            if sx != x:
                raise ValueError(f"{kind}_potential slabs do not tile x: expected StartX {x}, found {sx}")
            rec = np.memmap(f, dtype=_REC, mode="r", offset=16)
            if rec.size != nx * N * N:
                raise ValueError(f"{f}: {rec.size} records, expected {nx * N * N}")
            rec = rec.reshape(nx, N, N)
            i, j, k = np.ogrid[:nx, :N, :N]
            if not np.array_equal(rec["ind"], (i * N + j) * W + k):
                raise ValueError(f"{f}: record coordinates out of order")
            arr[sx:sx + nx] = rec["pot"]
            del rec
            x += nx
        if x != N:
            raise ValueError(f"{kind}_potential slabs cover {x} of {N} x-planes")
        if remove:
            for *_, f in slabs:
                os.remove(f)
        out[f"{kind}_potential"] = arr
    return out
