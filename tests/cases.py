"""Shared configurations and fixtures for the parity tests.

The configurations follow the reference's parameter-file tags
(Aarambam/LPT/Basis/read_param.c:27-184).  Data fixtures under tests/golden/
are the reference's shipped tables (Aarambam/defaults/).
"""
import os

import numpy as np

from aarambam_amd import workloads as _w

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

UNITS = _w.UNITS


def base_config(**kw):
    """Plain FP64 CPU-runnable case: Nmesh=64, Lbox=1000, EH P(k), local template."""
    return _w.config(**kw)


def load_transfer():
    d = np.loadtxt(os.path.join(GOLDEN, "TransferFunctionConverted.dat"))
    return d[:, 0].copy(), d[:, 1].copy()


def alpha_rows_file(path=None):
    path = path or os.path.join(GOLDEN, "AlphaTable.dat")
    return np.loadtxt(path).reshape(-1, 4)


def x0_file(Nm, path=None):
    path = path or os.path.join(GOLDEN, "AiTable.dat")
    d = np.loadtxt(path).reshape(-1, 3)
    x0 = np.zeros((Nm, 2))
    for r in d:
        n = int(r[0])
        if n < Nm:
            x0[n] = r[1:]
    return x0


local_rows = _w.local_rows  # SingleField.py:19-27
equilateral_rows = _w.equilateral_rows  # SingleField.py:45-78
le_rows = _w.le_rows
dense12_rows = _w.dense12_rows
lattice_glass = _w.lattice_glass


def x0_closed_form(Nm, ns, Box, Nmax):
    """AiTable x0[p] = P_p(-1 - 2 kmin^s/(kmax^s - kmin^s)), decompose.py:176-216."""
    return _w.x0_table(Nm, ns, Box, Nmax)


def job_inputs(cfg, rows=None, x0=None, glass_n=None):
    """(alpha table, x0 table, glass positions, glass box) for a config."""
    from oracle import pyoracle as orc
    Nm = cfg["N_modes"]
    rows = local_rows() if rows is None else rows
    alpha = orc.alpha_table(Nm, rows)
    if x0 is None:
        x0 = x0_closed_form(Nm, cfg["PrimordialIndex"], cfg["Box"], cfg["Nmax"])
    n = glass_n or cfg["Nmesh"] // cfg["GlassTileFac"]
    return alpha, x0, lattice_glass(n, 100.0), 100.0


def shipped_table_config(N=32, **kw):
    """The reference's shipped 15-mode ScalarI tables (ns=0.9660499, Nmax=256)."""
    cfg = base_config(Nmesh=N, Nsample=N, Nmax=256, N_modes=15, PrimordialIndex=0.9660499, **kw)
    return cfg, alpha_rows_file(), x0_file(15)


PARAM_TEMPLATE = """
Nmesh                               %(Nmesh)d
Nsample                             %(Nsample)d
Nmax                                %(Nmax)d
Box                                 %(Box)f
FileBase                            %(FileBase)s
OutputDir                           %(OutputDir)s
GlassFile                           %(GlassFile)s
GlassTileFac                        %(GlassTileFac)d

Seed                                %(Seed)d

%% a comment line
Omega                               %(Omega)f
OmegaLambda                         %(OmegaLambda)f
OmegaBaryon                         %(OmegaBaryon)f #trailing comment
OmegaDM_2ndSpecies                  0.00
HubbleParam                         %(HubbleParam)f
Sigma8                              %(Sigma8)f
PrimordialIndex                     %(PrimordialIndex)f
Redshift                            %(Redshift)f
Fnl                                 %(Fnl)f
N_modes                             %(N_modes)d

FixedAmplitude                      %(FixedAmplitude)d
PhaseFlip                           %(PhaseFlip)d
SphereMode                          %(SphereMode)d
SavePotentialField                  %(SavePotentialField)d

WhichSpectrum                       %(WhichSpectrum)d
WhichTransfer                       %(WhichTransfer)d

FileWithInputSpectrum               ThisIsAFakeStringThatIsNeverUsed
FileWithInputTransfer               %(TransferPath)s
FileWithAlphaCoeff                  %(AlphaPath)s
FileWithX0Coeff                     %(AiPath)s

NumFilesWrittenInParallel           %(NumFilesWrittenInParallel)d

InputSpectrum_UnitLength_in_cm      3.085678e24
UnitLength_in_cm                    3.085678e24
UnitMass_in_g                       1.989e43
UnitVelocity_in_cm_per_s            1e5

ShapeGamma                          0.0
WDM_On                              0
WDM_Vtherm_On                       0
WDM_PartMass_in_kev                 10.0
"""


# the three extra tags of the ResBasis parameter file (Aarambam/utils.py:72-78)
RES_TEMPLATE = """
A_res               %(A_res)f
w_res               %(w_res)f
phi_res             %(phi_res)f
"""


def write_param_file(path, cfg, res=False, **files):
    d = dict(cfg)
    d.setdefault("FileBase", "ics")
    d.setdefault("OutputDir", os.path.dirname(path))
    d.setdefault("SavePotentialField", 0)
    d.setdefault("NumFilesWrittenInParallel", 1)
    d.update(files)
    with open(path, "w") as f:
        f.write(PARAM_TEMPLATE % d + (RES_TEMPLATE % d if res else ""))


def res_config(which, A=0.5, w=40.0, phi=0.3, **kw):
    """ResBasis case: WhichSpectrum 0/1/2 with the oscillation A cos(w k + phi) or
    A cos(w ln k + phi) (Aarambam/LPT/ResBasis/power.c:240-260)."""
    return base_config(Variant=1, WhichSpectrum=which, A_res=A, w_res=w, phi_res=phi, **kw)


def write_lattice_glass(path, n=64, box=100.0):
    """Gadget-1 glass file in the layout of Aarambam/defaults/dummy_glass_dmonly_64.dat:
    256-byte header (type-2 particles, BoxSize, num_files=1), then the position block."""
    import struct
    pos = lattice_glass(n, box)
    hdr = bytearray(256)
    npart = [0, 0, n ** 3, 0, 0, 0]
    hdr[0:24] = struct.pack("6I", *npart)
    hdr[96:120] = struct.pack("6I", *npart)
    hdr[124:128] = struct.pack("i", 1)
    hdr[128:136] = struct.pack("d", box)
    with open(path, "wb") as f:
        f.write(struct.pack("i", 256))
        f.write(bytes(hdr))
        f.write(struct.pack("i", 256))
        f.write(struct.pack("i", 12 * n ** 3))
        f.write(pos.tobytes())
        f.write(struct.pack("i", 12 * n ** 3))
    return pos


# sha256 of Aarambam/defaults/dummy_glass_dmonly_64.dat
REFERENCE_GLASS_SHA256 = "684dc7e1ce49983c3948cca40abc320726737090469df96c55574bdc7c6f4579"


def scalari16_rows():
    """AlphaTable rows of ScalarI (mass 1) decomposed into 16 basis modes (n_s 0.9649, L 2000,
    Nmax 512, ModeTol 1e-4, MaxModeCount 150), written by the reference's own decomposition
    (tests/golden/make_expansion_vectors.py)."""
    return np.loadtxt(os.path.join(GOLDEN, "scalari16_alphatable.txt"), comments="#").reshape(-1, 4)


def example_config(N=512, Nsample=None, Nmax=None, Box=1000.0, **kw):
    """The reference's example run (examples/MakeICs.py:11-25, 52-60 with utils.py:9-58): the
    shipped 15-mode ScalarI tables, tabulated CAMB transfer, OmegaBaryon 0, cube cut at
    Nsample = Nmesh / 2 with Nmax = 256, a 64^3 glass tiled 4 times (256^3 particles on a
    512^3 mesh), seed 42.  Smaller N scale Nsample, Nmax and the box with it (the glass is
    then an (N/8)^3 lattice)."""
    tk, tt = load_transfer()
    f = N / 512
    cfg = base_config(Nmesh=N, Nsample=Nsample or N // 2, Nmax=Nmax or int(256 * f), N_modes=15,
                      Box=Box * f, GlassTileFac=4, Seed=42, Omega=0.3175, OmegaLambda=1 - 0.3175,
                      OmegaBaryon=0.0, HubbleParam=0.6711, Sigma8=0.834, PrimordialIndex=0.9660499,
                      Redshift=127.0, Fnl=100.0, WhichTransfer=2, transfer_k=tk, transfer_t=tt)
    cfg.update(kw)
    return cfg
