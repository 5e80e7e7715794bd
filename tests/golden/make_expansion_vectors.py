"""Writes the basis-expansion vectors in this directory from the reference's own
decomposition code -- run here only, in a tree that holds the reference.

The OMP loop of Basis.basis_expansion (Aarambam/basis/cmbbest.pyx:320-490) picks
modes by the ten largest residual correlations and the smallest rise of a
condition number: no closed form fixes which modes it adds, so these vectors
are the only pin of that order.  The reference module is built from its sources
by oracle/build_ref_cmbbest.py into oracle/_ref/ (`make -C oracle ref_cmbbest`); this
script imports it with the reference's decompose.py and the ScalarI model
(Aarambam/models/Colliders.py:4-72) and writes:

* expansion_grid40.npz: the Basis grid for 16 modes on a 40-point k grid
  (k, w, tetrapyd indices, mode functions, 3-D mode indices), the ScalarI
  (mass 1) and equilateral shapes on it, and basis_expansion's alpha,
  correlation and MSE for both with decompose.py's settings (ScalarI: ModeTol
  1e-4, MaxModeCount 150; equilateral: ModeTol 1e-6, MaxModeCount 60);
* scalari16_alphatable.txt: the AlphaTable rows with nonzero alpha that
  BasicBasisDecompose.generate_alpha (decompose.py:84-174, 150-point k grid)
  makes for ScalarI (mass 1) with 16 modes, n_s 0.9649, L 2000, Nmax 512,
  ModeTol 1e-4, MaxModeCount 150, and its correlation.

usage: python tests/golden/make_expansion_vectors.py [reference root]
"""
import contextlib
import importlib
import importlib.util
import io
import os
import re
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
NS, LBOX, NMAX, NMODES = 0.9649, 2000.0, 512, 16


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    # the built module is the top-level `cmbbest` (its pickled functions name it so in the worker
    # processes of the expansion's joblib pool, which inherit PYTHONPATH); decompose.py's
    # `from . import cmbbest` resolves through a package that holds it
    lib = os.path.join(ROOT, "oracle", "_ref")
    sys.path.insert(0, lib)
    os.environ["PYTHONPATH"] = lib + os.pathsep + os.environ.get("PYTHONPATH", "")
    best = importlib.import_module("cmbbest")
    pkg = types.ModuleType("refbasis")
    pkg.__path__ = [lib]
    sys.modules["refbasis"] = pkg
    sys.modules["refbasis.cmbbest"] = best
    dec = load("refbasis.decompose", os.path.join(REF, "Aarambam", "basis", "decompose.py"))
    coll = load("refcolliders", os.path.join(REF, "Aarambam", "models", "Colliders.py"))

    class ScalarIDecompose(coll.ScalarI, dec.BasicBasisDecompose):
        pass

    def decomposer(**kw):
        d = ScalarIDecompose(N_modes=NMODES, n_s=NS, Lbox=LBOX, Nmax=NMAX, ModeTol=1e-4, MaxModeCount=150, mass=1.0)
        d.args.update(kw)
        return d

    kmin, kmax = 2 * np.pi / LBOX, 2 * np.pi / LBOX * NMAX
    basis = best.Basis(mode_p_max=NMODES, parameter_n_scalar=NS, k_grid_size=40, k_min=kmin, k_max=kmax)
    p1, p2, p3 = basis.mode_indices
    msk = (p1 < 4) & np.invert(((p1 == 1) & (p2 == 1) & (p3 == 1)) | ((p1 == 2) & (p2 == 1) & (p3 == 0))
                               | ((p1 == 3) & (p2 == 0) & (p3 == 0)))  # decompose.py:125-129
    k1, k2, k3 = basis.tetrapyd_grid
    out = dict(k=basis.k_grid, w=basis.tetrapyd_grid_weights, i1=basis.tetrapyd_indices[0],
               i2=basis.tetrapyd_indices[1], i3=basis.tetrapyd_indices[2], evals=basis.mode_function_evaluations,
               p1=p1, p2=p2, p3=p3, bad_modes=msk, params=np.array([NS, kmin, kmax]))
    shapes = {"scalari": (decomposer().shape_function, 1e-4, 150),
              "equil": (best.Model("equilateral").shape_function, 1e-6, 60)}
    for name, (fn, tol, mm) in shapes.items():
        model = best.Model(shape_type="custom", shape_name=name, shape_function=fn)
        alpha, cov, corr, mse = basis.basis_expansion([model], rtol=1e-10, Niter=1e3, bad_modes=msk, RR_tol=tol,
                                                      max_modes=mm)
        out[f"S_{name}"] = fn(k1, k2, k3)
        out[f"alpha_{name}"] = alpha[0]
        out[f"corr_{name}"] = np.array([corr[0], mse[0], tol, mm])
        print(name, "selected", int((alpha[0] != 0).sum()), "correlation", corr[0], flush=True)
    np.savez_compressed(os.path.join(HERE, "expansion_grid40.npz"), **out)

    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):  # generate_alpha prints its correlation
        res = decomposer().generate_alpha(verbose=True)
    corr = float(re.search(r"correlation : ([0-9.]+)", buf.getvalue()).group(1))
    t = res["AlphaTable"]
    t = t[t[:, 3] != 0]
    np.savetxt(os.path.join(HERE, "scalari16_alphatable.txt"), t, fmt=["%d", "%d", "%d", "%.17g"],
               header=f"ScalarI mass 1, N_modes {NMODES}, n_s {NS}, Lbox {LBOX}, Nmax {NMAX}, ModeTol 1e-4, "
                      f"MaxModeCount 150; correlation {corr:.10f}; p1 p2 p3 alpha (nonzero rows)")
    print("scalari16 rows", len(t), flush=True)


if __name__ == "__main__":
    main()
