// Fused z kernels compiled at run time for one job's product schedule (hiprtc).
//
// The library's own fused z kernel (k_z_ng) reads the NG sum's product schedule -- the greedy
// vertex cover of each output group, as 16-byte records -- at run time, so every read of a mode
// field is a run of indexed register moves (one VALU move per value) and every block of four
// records a scalar load.  For the schedule of one job the factors, parts, slots and coefficients
// are constants: the kernel compiled here writes them out (the same fused multiply-adds in the
// same order, so the results are bit-identical to k_z_ng) and reads every mode slot with a
// constant index (the mode loop stays a loop: unrolled, it needs more than 256 registers).  The body is z_ng_body (aar_zng.h), whose headers are
// embedded in the library at build time.  Kernels are cached per device by their source, so a
// stream of jobs with one table compiles once.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "aar_jit.h"

namespace {

// kJitHeaderNames / kJitHeaderText: aar_device.h, aar_fft.h and aar_zng.h, written by the build
#include "aar_jit_headers.inc"

struct Entry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  int status = AAR_OK;
  std::string log;
};
std::mutex g_mu;
std::map<std::pair<int, std::string>, Entry> g_cache;

std::string coef_literal(int fp64, const NgRec& r) {
  char b[64];
  if (fp64) {
    unsigned long long bits;
    memcpy(&bits, &r.c, 8);
    snprintf(b, sizeof b, "__longlong_as_double(0x%016llxLL)", bits);
  } else {
    snprintf(b, sizeof b, "__int_as_float(0x%08x)", (unsigned)r.aux);
  }
  return b;
}

}  // namespace

void jit_zng_shape(int fp64, int H, int M, int* rows, int* threads) {
  // ZNGCfg<T, H, M, 0> (aar_zng.h); the generated source checks that they agree
  const int E = fp64 ? 2 : ((H == 512 && M <= 6) ? 8 : 4);
  const int C = (!fp64 && (H < 512 || E == 8)) ? 4 : 2;
  *rows = C;
  *threads = C * H / E;
}

std::string jit_zng_source(int fp64, int H, int M, int K, const int* groups, const NgRec4* recs) {
  int C, NT;
  jit_zng_shape(fp64, H, M, &C, &NT);
  const char* T = fp64 ? "double" : "float";
  std::string s;
  char b[512];
  s += "#include \"aar_zng.h\"\n";
  snprintf(b, sizeof b,
           "using T = %s;\nusing C2 = cx_t<T>;\nconstexpr int kH = %d, kM = %d;\n"
           "static_assert(ZNGCfg<T, kH, kM, 0>::C == %d && ZNGCfg<T, kH, kM, 0>::NT == %d, \"launch shape\");\n",
           T, H, M, C, NT);
  s += b;
  s += "struct Products {\n  template <int E, class Get>\n"
       "  __device__ __forceinline__ void operator()(int kk, Get& get, C2* hv) const {\n"
       "    C2 u[E], b[E];\n    switch (kk) {\n";
  auto rec = [&](int q) -> const NgRec& { return recs[q / 4].q[q % 4]; };
  for (int kk = 0; kk < K; ++kk) {
    snprintf(b, sizeof b, "      case %d: {\n", kk);
    s += b;
    for (int r = groups[kk]; r < groups[kk + 1];) {
      const int v = rec(r).slot, n = rec(r).aux;
      s += "#pragma unroll\n        for (int e = 0; e < E; ++e) u[e] = cmake<C2>(0, 0);\n";
      for (int t = 1; t <= n; ++t) {
        const NgRec& p = rec(r + t);
        snprintf(b, sizeof b,
                 "        get(%d, b);\n        {\n          const T cf = %s;\n#pragma unroll\n"
                 "          for (int e = 0; e < E; ++e) {\n            u[e].x = fma(cf, b[e].x, u[e].x);\n"
                 "            u[e].y = fma(cf, b[e].y, u[e].y);\n          }\n        }\n",
                 p.slot, coef_literal(fp64, p).c_str());
        s += b;
      }
      snprintf(b, sizeof b,
               "        get(%d, b);\n#pragma unroll\n        for (int e = 0; e < E; ++e) {\n"
               "          hv[e].x = fma(b[e].x, u[e].x, hv[e].x);\n          hv[e].y = fma(b[e].y, u[e].y, hv[e].y);\n"
               "        }\n",
               v);
      s += b;
      r += (n + 4) & ~3;
    }
    s += "      } break;\n";
  }
  s += "      default: break;\n    }\n  }\n};\n";
  s += "extern \"C\" __global__ void __launch_bounds__((ZNGCfg<T, kH, kM, 0>::NT))\n"
       "__attribute__((amdgpu_waves_per_eu(ZNGCfg<T, kH, kM, 0>::WPE)))\n"
       "    aar_zng_jit(C2* base, int M, int K, Geom g, const C2* __restrict__ twg) {\n"
       "  z_ng_run<T, kH, kM, 0>(base, M, K, g, twg, Products{});\n}\n";
  return s;
}

int jit_zng_kernel(int device, const std::string& src, JitZng* out, std::string* log) {
  std::lock_guard<std::mutex> lk(g_mu);
  Entry& e = g_cache[{device, src}];
  if (!e.fn && e.status == AAR_OK) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "aar_zng_jit.hip", kJitHeaderCount, kJitHeaderText,
                            kJitHeaderNames) != HIPRTC_SUCCESS) {
      e.status = AAR_ERR_HIP;
    } else {
      // the library's own flags: no contraction into fused multiply-adds, no packed FP32
      const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
      const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
      size_t n = 0;
      if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
        e.log.resize(n);
        hiprtcGetProgramLog(prog, &e.log[0]);
      }
      std::vector<char> code;
      if (cr == HIPRTC_SUCCESS && hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS) {
        code.resize(n);
        if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
      }
      hiprtcDestroyProgram(&prog);
      if (code.empty()) {
        e.status = AAR_ERR_UNSUPPORTED;
      } else if (hipSetDevice(device) != hipSuccess || hipModuleLoadData(&e.mod, code.data()) != hipSuccess ||
                 hipModuleGetFunction(&e.fn, e.mod, "aar_zng_jit") != hipSuccess) {
        e.fn = nullptr;
        e.status = AAR_ERR_HIP;
      }
    }
  }
  if (log) *log = e.log;
  out->fn = e.fn;
  return e.fn ? AAR_OK : e.status;
}

int jit_zng_launch(const JitZng& k, void* base, int M, int K, const Geom& g, const void* tw, hipStream_t st) {
  Geom gg = g;
  void* b = base;
  const void* t = tw;
  void* args[] = {&b, &M, &K, &gg, &t};
  const unsigned blocks = (unsigned)((long long)g.N * g.N / k.rows);
  return hipModuleLaunchKernel(k.fn, blocks, 1, 1, (unsigned)k.threads, 1, 1, 0, st, args, nullptr) == hipSuccess
             ? AAR_OK
             : AAR_ERR_HIP;
}
