// libismc_diag: measurement probes behind the transport kernel's roofline
// and the byte counters of its profiles.  Not part of libismc (no
// counterpart in the reference); declarations in include/ismc_diag.h.
//
// 1. VALU issue costs of single instruction kinds on the device (gfx950).
//
// Each kernel runs one instruction kind in 8 independent dependency chains
// per lane (so issue, not latency, sets the pace at 8 waves per SIMD), over
// a grid of 64 workgroups of 256 lanes per CU (so no SIMD sits idle while
// others finish).  The first lane of every workgroup reads the shader
// clock counter (s_memtime) and the 100 MHz real-time counter
// (s_memrealtime) around its loop; their ratio is the clock the loop ran
// at.  cycles per wave64 instruction = SIMDs x clock / (instructions / s).
//
// This is the measurement behind the roofline of the transport kernel: a
// full-rate FP32 operation (add, FMA, compare, logic) takes about 2 cycles
// for a 64-lane wave, i.e. 32 lanes per SIMD per cycle without packed
// math; 32-bit integer multiplies, mad_u64_u32 and FP64 adds take twice
// that, and transcendentals (rcp, sqrt, log, exp, sin) four times.  The
// three-input v_bitop3_b32 (Philox's key XORs) is measured as well.
//
// 2. A plain streaming copy of known size: a profile run reads what the
// hardware byte counters (FETCH_SIZE / WRITE_SIZE) give for it, and scales
// the transport kernel's counts by the same factor.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/ismc_diag.h"

namespace issue {

constexpr int kChains = 8;
constexpr int kBlocksPerCU = 64;
constexpr int kIters = 2048;
constexpr int kLaunches = 5;

enum Op : int {
  ADD_F32, FMA_F32, SHIFT_XOR, MUL_LO_U32, MAD_U64_U32, ADD_F64, RCP_F32, SQRT_F32, LOG_F32,
  EXP_F32, SIN_F32, BITOP3, RCP_ADD, MAD64_ADD, BITOP3_ADD, FMA_VGPR, MUL_LITERAL, N_OPS
};

// instruction name and wave instructions per chain step
struct OpInfo { const char* name; int insts; };
constexpr OpInfo kOps[N_OPS] = {
    {"v_add_f32", 1},     {"v_fma_f32", 1},     {"v_lshrrev_b32+v_xor_b32", 2},
    {"v_mul_lo_u32", 1},  {"v_mad_u64_u32", 1}, {"v_add_f64", 1},
    {"v_rcp_f32", 1},     {"v_sqrt_f32", 1},    {"v_log_f32", 1},
    {"v_exp_f32", 1},     {"v_sin_f32", 1},     {"v_bitop3_b32", 1},
    // pairs of independent chains: whether the slower kind overlaps the
    // full-rate add (then a pair costs about the slower one alone)
    {"v_rcp_f32|v_add_f32", 2}, {"v_mad_u64_u32|v_add_f32", 2}, {"v_bitop3_b32|v_add_f32", 2},
    // FMA with every operand in a VGPR, and a multiply by a 32-bit literal:
    // which of the two makes the literal-addend FMA above slower
    {"v_fma_f32 (VGPR operands)", 1}, {"v_mul_f32 (literal)", 1},
};

// out: one word per lane; clk: two counts per workgroup.  The grid is
// exactly blocks x 256 lanes, the buffers are sized for it.
template <int OP>
__global__ __launch_bounds__(256) void chain_kernel(uint32_t* out, unsigned long long* clk, int iters,
                                                    uint32_t seed) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  float f[kChains], f2[kChains];
  uint32_t u[kChains];
  uint64_t w[kChains];
  double d[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    f[c] = 1.0f + (float)(t + c) * 1e-7f;
    f2[c] = 2.0f + (float)(t + c) * 1e-7f;
    u[c] = t * 2654435761u + (uint32_t)c + seed;
    w[c] = ((uint64_t)u[c] << 32) | (t + c);
    d[c] = 1.0 + (double)(t + c) * 1e-9;
  }
  const float m = 0.999f + (float)seed * 1e-9f, a = 1e-4f;
  // per-lane values: kept in VGPRs
  const float mv = 0.999f + (float)t * 1e-12f, av = 1e-4f + (float)t * 1e-13f;
  const double da = 1e-12 * seed;
  const unsigned long long c0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if constexpr (OP == ADD_F32) f[c] = f[c] + a;
      if constexpr (OP == FMA_F32) f[c] = fmaf(f[c], m, a);
      if constexpr (OP == SHIFT_XOR) u[c] = u[c] ^ (u[c] >> 3);
      if constexpr (OP == MUL_LO_U32) u[c] = u[c] * 0xCD9E8D57u;
      if constexpr (OP == MAD_U64_U32) w[c] = (uint64_t)(uint32_t)w[c] * 0xD2511F53u + w[c];
      if constexpr (OP == ADD_F64) d[c] = d[c] + da;
      if constexpr (OP == RCP_F32) f[c] = __builtin_amdgcn_rcpf(f[c]);
      if constexpr (OP == SQRT_F32) f[c] = __builtin_amdgcn_sqrtf(f[c]);
      if constexpr (OP == LOG_F32) f[c] = __builtin_amdgcn_logf(f[c]);
      if constexpr (OP == EXP_F32) f[c] = __builtin_amdgcn_exp2f(f[c]);
      if constexpr (OP == SIN_F32) f[c] = __builtin_amdgcn_sinf(f[c]);
      if constexpr (OP == BITOP3) u[c] = __builtin_amdgcn_bitop3_b32(u[c], t, seed, 0x96);
      if constexpr (OP == RCP_ADD) { f[c] = __builtin_amdgcn_rcpf(f[c]); f2[c] = f2[c] + a; }
      if constexpr (OP == MAD64_ADD) {
        w[c] = (uint64_t)(uint32_t)w[c] * 0xD2511F53u + w[c];
        f2[c] = f2[c] + a;
      }
      if constexpr (OP == BITOP3_ADD) { u[c] = __builtin_amdgcn_bitop3_b32(u[c], t, seed, 0x96); f2[c] = f2[c] + a; }
      if constexpr (OP == FMA_VGPR) f[c] = fmaf(f[c], mv, av);
      if constexpr (OP == MUL_LITERAL) f[c] = f[c] * 0.99990004f;
    }
  }
  const unsigned long long c1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c)
    r ^= __float_as_uint(f[c]) ^ __float_as_uint(f2[c]) ^ u[c] ^ (uint32_t)w[c] ^ (uint32_t)(w[c] >> 32) ^
         (uint32_t)__double2loint(d[c]);
  out[t] = r;
  if (threadIdx.x == 0) {
    clk[2 * blockIdx.x] = c1 - c0;
    clk[2 * blockIdx.x + 1] = r1 - r0;
  }
}

__global__ void copy_probe_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

using KernelFn = void (*)(uint32_t*, unsigned long long*, int, uint32_t);

template <int... I>
constexpr auto kernel_table(std::integer_sequence<int, I...>) {
  return std::array<KernelFn, sizeof...(I)>{&chain_kernel<I>...};
}

thread_local std::string g_err;

int fail(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  g_err = buf;
  return code;
}

}  // namespace issue

extern "C" {

int32_t ismc_issue_rate_ops(void) { return issue::N_OPS; }

const char* ismc_issue_rate_name(int32_t op) {
  return (op >= 0 && op < issue::N_OPS) ? issue::kOps[op].name : nullptr;
}

const char* ismc_diag_last_error(void) { return issue::g_err.c_str(); }

int ismc_issue_rate(int device, int32_t op, double* wave_insts_per_s, double* shader_ghz) {
  using namespace issue;
  if (op < 0 || op >= N_OPS || !wave_insts_per_s || !shader_ghz) {
    g_err = "ismc_issue_rate: op outside 0..ismc_issue_rate_ops()-1 or null output";
    return ISMC_DIAG_EINVAL;
  }
  static const auto table = kernel_table(std::make_integer_sequence<int, N_OPS>{});
  const KernelFn kern = table[op];
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(ISMC_DIAG_EDEVICE, "hipSetDevice", e);
  int n_cu = 0;
  if ((e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
    return fail(ISMC_DIAG_EDEVICE, "hipDeviceGetAttribute", e);
  const int blocks = n_cu * kBlocksPerCU;
  uint32_t* out = nullptr;
  unsigned long long* clk = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  std::vector<unsigned long long> h(2 * (size_t)blocks);
  float ms = 0.f;
  double cyc = 0.0, rt = 0.0;
  if ((e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&clk, h.size() * sizeof(unsigned long long))) != hipSuccess) {
    rc = fail(ISMC_DIAG_ENOMEM, "hipMalloc", e);
    goto done;
  }
  if ((e = hipStreamCreate(&s)) != hipSuccess || (e = hipEventCreate(&e0)) != hipSuccess ||
      (e = hipEventCreate(&e1)) != hipSuccess) {
    rc = fail(ISMC_DIAG_EDEVICE, "hipStreamCreate/hipEventCreate", e);
    goto done;
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, out, clk, kIters, 1u);  // warm-up
  if ((e = hipEventRecord(e0, s)) != hipSuccess) { rc = fail(ISMC_DIAG_EDEVICE, "hipEventRecord", e); goto done; }
  for (int r = 0; r < kLaunches; ++r)
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, out, clk, kIters, (uint32_t)(r + 2));
  if ((e = hipGetLastError()) != hipSuccess) { rc = fail(ISMC_DIAG_EDEVICE, "chain_kernel", e); goto done; }
  if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
      (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
    rc = fail(ISMC_DIAG_EDEVICE, "hipEvent timing", e);
    goto done;
  }
  // clock of the last launch's loops
  if ((e = hipMemcpy(h.data(), clk, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost)) !=
      hipSuccess) {
    rc = fail(ISMC_DIAG_EDEVICE, "hipMemcpy", e);
    goto done;
  }
  for (int b = 0; b < blocks; ++b) { cyc += (double)h[2 * b]; rt += (double)h[2 * b + 1]; }
  *wave_insts_per_s = (double)kLaunches * blocks * 4.0 * kIters * kChains * kOps[op].insts / (ms * 1e-3);
  *shader_ghz = rt > 0 ? cyc / rt * 0.1 : 0.0;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (out) (void)hipFree(out);
  if (clk) (void)hipFree(clk);
  return rc;
}

int ismc_copy_probe(int device, int64_t bytes, float* kernel_ms) {
  using namespace issue;
  if (bytes <= 0 || (bytes % 16) != 0 || !kernel_ms) {
    g_err = "ismc_copy_probe: needs a positive multiple of 16 bytes and an output";
    return ISMC_DIAG_EINVAL;
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(ISMC_DIAG_EDEVICE, "hipSetDevice", e);
  int n_cu = 0;
  if ((e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
    return fail(ISMC_DIAG_EDEVICE, "hipDeviceGetAttribute", e);
  void* a = nullptr;
  void* b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  const size_t n = (size_t)bytes / 16;
  const int grid = n_cu * 8;
  if ((e = hipMalloc(&a, (size_t)bytes)) != hipSuccess || (e = hipMalloc(&b, (size_t)bytes)) != hipSuccess) {
    rc = fail(ISMC_DIAG_ENOMEM, "hipMalloc", e);
    goto done;
  }
  if ((e = hipStreamCreate(&s)) != hipSuccess || (e = hipEventCreate(&e0)) != hipSuccess ||
      (e = hipEventCreate(&e1)) != hipSuccess) {
    rc = fail(ISMC_DIAG_EDEVICE, "hipStreamCreate/hipEventCreate", e);
    goto done;
  }
  hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, (const uint4*)b, (uint4*)a, n);  // warm
  if ((e = hipEventRecord(e0, s)) != hipSuccess) { rc = fail(ISMC_DIAG_EDEVICE, "hipEventRecord", e); goto done; }
  hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, (const uint4*)a, (uint4*)b, n);
  if ((e = hipGetLastError()) != hipSuccess) { rc = fail(ISMC_DIAG_EDEVICE, "copy_probe_kernel", e); goto done; }
  if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
      (e = hipEventElapsedTime(kernel_ms, e0, e1)) != hipSuccess) {
    rc = fail(ISMC_DIAG_EDEVICE, "hipEvent timing", e);
    goto done;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
