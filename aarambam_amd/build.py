"""Build the HIP library (libaarambam_hip.so) and the 2LPTBasis / 2LPTResBasis executables
in-tree for gfx950.

Usage: python -m aarambam_amd.build   (or aarambam_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaarambam_hip.so")
EXE = os.path.join(HERE, "bin", "2LPTBasis")
EXE_RES = os.path.join(HERE, "bin", "2LPTResBasis")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["aar_kernels.hip", "aar_stats.hip", "aar_modal.hip", "aar_job.cpp", "aar_host.cpp", "aar_serve.cpp",
           "aar_modal_host.cpp", "aar_jit.cpp"]
# headers the run-time compiled fused z kernels include (aar_jit.cpp), embedded as string literals
JIT_HEADERS = ["aar_device.h", "aar_fft.h", "aar_zng.h"]
# -ffp-contract=off keeps the reference's operation order (no fused multiply-add
# contraction); packed FP32 instructions are disabled by target feature.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
         "-pthread"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def write_jit_headers():
    """_build/aar_jit_headers.inc: the headers of the run-time compiled kernels as raw string
    literals (kJitHeaderNames, kJitHeaderText, kJitHeaderCount)."""
    names, texts = [], []
    for h in JIT_HEADERS:
        text = open(os.path.join(CSRC, h)).read()
        if ')AARJIT"' in text:
            raise RuntimeError(f"{h} holds the literal delimiter")
        names.append(f'"{h}"')
        texts.append('R"AARJIT(' + text + ')AARJIT"')
    out = (f"static const int kJitHeaderCount = {len(JIT_HEADERS)};\n"
           f"static const char* kJitHeaderNames[] = {{{', '.join(names)}}};\n"
           f"static const char* kJitHeaderText[] = {{\n" + ",\n".join(texts) + "};\n")
    path = os.path.join(BUILD, "aar_jit_headers.inc")
    with open(path + ".tmp", "w") as f:
        f.write(out)
    os.replace(path + ".tmp", path)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    write_jit_headers()

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + FLAGS + ["-I" + BUILD, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + ["-lhiprtc"])
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "aar_main.cpp")
    if os.path.exists(main_src):
        for exe, defs in [(EXE, []), (EXE_RES, ["-DAAR_RESBASIS"])]:
            _run([HIPCC] + FLAGS + defs + ["-x", "hip", main_src, "-o", exe, "-L" + HERE, "-laarambam_hip",
                                           "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
