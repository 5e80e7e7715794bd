"""Builds the reference's decomposition module (Aarambam/basis/cmbbest.pyx with
tetrapyd.c and arrays.c) from its sources where they lie into oracle/_ref/ --
TEST INFRASTRUCTURE ONLY.  Nothing under the product imports it: only the
script under tests/golden/ that writes the basis-expansion vectors loads it,
in a tree that holds the reference.

usage: python oracle/build_ref_cmbbest.py <reference basis dir> <output dir>
"""
import os
import sys

import numpy
from Cython.Build import cythonize
from setuptools import Extension
from setuptools.dist import Distribution


def main(ref, out):
    os.makedirs(out, exist_ok=True)
    ext = Extension("cmbbest", sources=[os.path.join(ref, f) for f in ("cmbbest.pyx", "tetrapyd.c", "arrays.c")],
                    include_dirs=[numpy.get_include(), ref], extra_compile_args=["-O2", "-fopenmp"],
                    extra_link_args=["-fopenmp"])
    exts = cythonize([ext], build_dir=os.path.join(out, "cy"), language_level=3, quiet=True)
    cmd = Distribution({"ext_modules": exts}).get_command_obj("build_ext")
    cmd.build_lib = out
    cmd.build_temp = os.path.join(out, "tmp")
    cmd.ensure_finalized()
    cmd.run()


if __name__ == "__main__":
    main(os.path.abspath(sys.argv[1]), os.path.abspath(sys.argv[2]))
