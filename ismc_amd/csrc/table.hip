// Lookup-table uncertainty map for MI355X (gfx950).
//
// The lookup tables map a measured (R_d, T_d) pair, both in percent on a
// uniform P x P grid over [lo, hi], to mu_a and mu_s' (scattered simulation
// points interpolated onto the grid by the host).  The uncertainty of a
// cell (i, j) is estimated the way Lookup/lookup.py:347-376 does it: draw M
// pairs R ~ N(R_i, Rstd[i][j]), T ~ N(T_j, Tstd[i][j]), look each pair up by
// nearest grid point (find_nearest, lookup.py:296-299: argmin of |grid - v|,
// first index on ties), and report 100 * std / centre of the looked-up
// mu_a and mu_s' values (population std; centre = the table value at the
// cell's own nearest grid point).  A cell whose Rstd or Tstd is NaN or
// negative gets NaN, as in the reference; a NaN table value in a sample
// makes the cell NaN (numpy's std propagates it).
//
// The reference scans the whole grid axis for every sample (O(P) per
// lookup) in a Python loop over P^2 cells x M samples.  Here one thread
// owns one cell: the nearest index comes from rounding on the uniform grid
// and a check of the two neighbours against the actual grid values (so the
// first-minimum rule of argmin is kept), and the table gathers hit an 8 MB
// table that stays in L2 / the last-level cache.
//
// Normals: Philox4x32-10 (key = seed, counter = (cell, sample pair, 0,
// 0x7ab1e)), each 32-bit word w -> u = (w + 0.5) 2^-32 in (0, 1), and
// Box-Muller z0 = sqrt(-2 ln u0) cos(2 pi u1), z1 = sqrt(-2 ln u0) sin(2 pi u1)
// for the R and T draws of one sample (words 2, 3 give the next sample).
// FP64 throughout, like numpy.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/ismc.h"
#include "philox.hpp"

namespace tablek {

constexpr uint32_t kStreamTag = 0x7ab1eu;

__device__ __forceinline__ double u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

// argmin_k |g[k] - v| over the uniform grid g[0..P-1] (first index on ties)
__device__ __forceinline__ int nearest(const double* __restrict__ g, int P, double lo, double step, double v) {
  double f = (v - lo) / step + 0.5;
  f = fmin(fmax(f, 0.0), (double)(P - 1));
  const int k0 = (int)f;
  int best = k0 > 0 ? k0 - 1 : k0;
  double bd = fabs(g[best] - v);
  const int kend = k0 + 1 < P ? k0 + 1 : P - 1;
  for (int k = best + 1; k <= kend; ++k) {
    const double d = fabs(g[k] - v);
    if (d < bd) { bd = d; best = k; }
  }
  return best;
}

struct Args {
  const double* mua;   // [P][P], index [i (R)][j (T)]
  const double* mus;
  const double* rstd;
  const double* tstd;
  const double* Rv;    // [P]
  const double* Tv;    // [P]
  double* mua_unc;
  double* mus_unc;
  int P, M;
  double lo, step;
  uint32_t k0, k1;
};

__global__ __launch_bounds__(256) void uncertainty_kernel(Args a) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = a.P;
  if (cell >= P * P) return;
  const int i = cell / P, j = cell - i * P;
  const double rs = a.rstd[cell], ts = a.tstd[cell];
  if (isnan(rs) || isnan(ts) || rs < 0.0 || ts < 0.0) {
    a.mua_unc[cell] = NAN;
    a.mus_unc[cell] = NAN;
    return;
  }
  const double R = a.Rv[i], T = a.Tv[j];
  const int ci = nearest(a.Rv, P, a.lo, a.step, R);
  const int cj = nearest(a.Tv, P, a.lo, a.step, T);
  const double ca = a.mua[ci * P + cj], cs = a.mus[ci * P + cj];
  // sums of deviations from the centre value (shifted, to keep the
  // one-pass variance well conditioned)
  double sa = 0.0, qa = 0.0, ss = 0.0, qs = 0.0;
  const int pairs = (a.M + 1) / 2;
  for (int p = 0; p < pairs; ++p) {
    uint32_t w[4];
    ismc_rng::philox4x32_10((uint32_t)cell, 0u, (uint32_t)p, kStreamTag, a.k0, a.k1, w);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (2 * p + h >= a.M) break;
      const double u0 = u01(w[2 * h]), u1 = u01(w[2 * h + 1]);
      const double r = sqrt(-2.0 * log(u0));
      const double ang = 6.283185307179586 * u1;
      const double Rs = R + rs * (r * cos(ang));
      const double Ts = T + ts * (r * sin(ang));
      const int ri = nearest(a.Rv, P, a.lo, a.step, Rs);
      const int tj = nearest(a.Tv, P, a.lo, a.step, Ts);
      const double va = a.mua[ri * P + tj] - ca, vs = a.mus[ri * P + tj] - cs;
      sa += va; qa += va * va;
      ss += vs; qs += vs * vs;
    }
  }
  const double m = (double)a.M;
  // clamp rounding below zero; a NaN (a sample on a NaN cell) must stay NaN,
  // which fmax would drop
  double var_a = qa / m - (sa / m) * (sa / m);
  double var_s = qs / m - (ss / m) * (ss / m);
  var_a = var_a < 0.0 ? 0.0 : var_a;
  var_s = var_s < 0.0 ? 0.0 : var_s;
  a.mua_unc[cell] = 100.0 * sqrt(var_a) / ca;
  a.mus_unc[cell] = 100.0 * sqrt(var_s) / cs;
}

// For measurement m, the index of the grid cell c minimizing
//   ((pR[c] - R[m]) / R[m])^2 + ((pT[c] - T[m]) / max(T[m], 1e-6))^2
// with numpy's argmin rules: the first index among equal minima, and a NaN
// cost counts as the minimum (the first NaN wins).  FP64, the operations in
// that order and nothing contracted into an fma, so the costs are the ones
// numpy computes bit for bit.  One workgroup per measurement; its lanes
// stride over the cells, then reduce (cost, index) pairs in LDS.
constexpr int kArgminBlock = 256;

__device__ __forceinline__ bool before(double v, int i, double bv, int bi) {
  const bool vn = isnan(v), bn = isnan(bv);
  if (vn != bn) return vn;
  if (!vn && v != bv) return v < bv;
  return i < bi;
}

__global__ __launch_bounds__(kArgminBlock) void grid_argmin_kernel(const float* __restrict__ pR,
                                                                   const float* __restrict__ pT, int G,
                                                                   const double* __restrict__ R,
                                                                   const double* __restrict__ T,
                                                                   int32_t* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double s_v[kArgminBlock];
  __shared__ int s_i[kArgminBlock];
  const int m = blockIdx.x;
  const double r = R[m], t = T[m], tf = t > 1e-6 ? t : 1e-6;
  double bv = 0.0;
  int bi = 0x7fffffff;
  for (int c = threadIdx.x; c < G; c += kArgminBlock) {
    const double a = ((double)pR[c] - r) / r;
    const double b = ((double)pT[c] - t) / tf;
    const double v = a * a + b * b;
    if (bi == 0x7fffffff || before(v, c, bv, bi)) { bv = v; bi = c; }
  }
  s_v[threadIdx.x] = bv;
  s_i[threadIdx.x] = bi;
  __syncthreads();
  for (int h = kArgminBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      const int j = threadIdx.x + h;
      if (s_i[j] != 0x7fffffff && (s_i[threadIdx.x] == 0x7fffffff ||
                                   before(s_v[j], s_i[j], s_v[threadIdx.x], s_i[threadIdx.x]))) {
        s_v[threadIdx.x] = s_v[j];
        s_i[threadIdx.x] = s_i[j];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = s_i[0];
}

thread_local std::string g_err;

int fail(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  g_err = buf;
  return code;
}

}  // namespace tablek

extern "C" {

const char* ismc_table_last_error(void) { return tablek::g_err.c_str(); }

int ismc_table_uncertainty(int device, const double* mua, const double* mus, const double* rstd,
                           const double* tstd, const double* Rvals, const double* Tvals, int32_t P,
                           int32_t M, uint64_t seed, double* mua_unc, double* mus_unc) {
  using namespace tablek;
  if (!mua || !mus || !rstd || !tstd || !Rvals || !Tvals || !mua_unc || !mus_unc || P < 2 ||
      P > 8192 || M < 1) {
    g_err = "ismc_table_uncertainty: null pointer, P outside 2..8192 or M < 1";
    return ISMC_EINVAL;
  }
  const double lo = Rvals[0], step = (Rvals[P - 1] - Rvals[0]) / (double)(P - 1);
  if (!(step > 0.0) || Tvals[0] != lo || Tvals[P - 1] != Rvals[P - 1]) {
    g_err = "ismc_table_uncertainty: R and T grids must be the same increasing uniform grid";
    return ISMC_EINVAL;
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(ISMC_EDEVICE, "hipSetDevice", e);
  const size_t cells = (size_t)P * P, tb = cells * sizeof(double), gb = (size_t)P * sizeof(double);
  double* d = nullptr;
  e = hipMalloc(&d, 6 * tb + 2 * gb);
  if (e != hipSuccess) return fail(ISMC_ENOMEM, "hipMalloc", e);
  Args a{};
  a.mua = d; a.mus = d + cells; a.rstd = d + 2 * cells; a.tstd = d + 3 * cells;
  a.mua_unc = d + 4 * cells; a.mus_unc = d + 5 * cells;
  a.Rv = d + 6 * cells; a.Tv = d + 6 * cells + P;
  a.P = P; a.M = M; a.lo = lo; a.step = step;
  a.k0 = (uint32_t)seed; a.k1 = (uint32_t)(seed >> 32);
  int rc = 0;
  hipStream_t s = nullptr;
  const void* src[6] = {mua, mus, rstd, tstd, Rvals, Tvals};
  void* dst[6] = {(void*)a.mua, (void*)a.mus, (void*)a.rstd, (void*)a.tstd, (void*)a.Rv, (void*)a.Tv};
  const size_t sz[6] = {tb, tb, tb, tb, gb, gb};
  if ((e = hipStreamCreate(&s)) != hipSuccess) { rc = fail(ISMC_EDEVICE, "hipStreamCreate", e); goto out; }
  for (int k = 0; k < 6; ++k)
    if ((e = hipMemcpyAsync(dst[k], src[k], sz[k], hipMemcpyHostToDevice, s)) != hipSuccess) {
      rc = fail(ISMC_EDEVICE, "hipMemcpyAsync", e);
      goto out;
    }
  uncertainty_kernel<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(a);
  if ((e = hipGetLastError()) != hipSuccess) { rc = fail(ISMC_EDEVICE, "uncertainty_kernel", e); goto out; }
  if ((e = hipMemcpyAsync(mua_unc, a.mua_unc, tb, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(mus_unc, a.mus_unc, tb, hipMemcpyDeviceToHost, s)) != hipSuccess) {
    rc = fail(ISMC_EDEVICE, "hipMemcpyAsync", e);
    goto out;
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) rc = fail(ISMC_EDEVICE, "hipStreamSynchronize", e);
out:
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(d);
  return rc;
}

int ismc_grid_argmin(int device, const float* pR, const float* pT, int32_t G, const double* R,
                     const double* T, int32_t M, int32_t* idx) {
  using namespace tablek;
  if (!pR || !pT || !R || !T || !idx || G < 1 || M < 0) {
    g_err = "ismc_grid_argmin: null pointer, G < 1 or M < 0";
    return ISMC_EINVAL;
  }
  if (M == 0) return 0;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(ISMC_EDEVICE, "hipSetDevice", e);
  const size_t gb = (size_t)G * sizeof(float), mb = (size_t)M * sizeof(double), ib = (size_t)M * sizeof(int32_t);
  char* d = nullptr;
  e = hipMalloc(&d, 2 * gb + 2 * mb + ib + 64);
  if (e != hipSuccess) return fail(ISMC_ENOMEM, "hipMalloc", e);
  // 8-byte aligned: the two double arrays first
  double* dR = (double*)d;
  double* dT = dR + M;
  float* dpR = (float*)(dT + M);
  float* dpT = dpR + G;
  int32_t* dI = (int32_t*)(dpT + G);
  int rc = 0;
  hipStream_t s = nullptr;
  const void* src[4] = {R, T, pR, pT};
  void* dst[4] = {dR, dT, dpR, dpT};
  const size_t sz[4] = {mb, mb, gb, gb};
  if ((e = hipStreamCreate(&s)) != hipSuccess) { rc = fail(ISMC_EDEVICE, "hipStreamCreate", e); goto out; }
  for (int k = 0; k < 4; ++k)
    if ((e = hipMemcpyAsync(dst[k], src[k], sz[k], hipMemcpyHostToDevice, s)) != hipSuccess) {
      rc = fail(ISMC_EDEVICE, "hipMemcpyAsync", e);
      goto out;
    }
  grid_argmin_kernel<<<(unsigned)M, kArgminBlock, 0, s>>>(dpR, dpT, G, dR, dT, dI);
  if ((e = hipGetLastError()) != hipSuccess) { rc = fail(ISMC_EDEVICE, "grid_argmin_kernel", e); goto out; }
  if ((e = hipMemcpyAsync(idx, dI, ib, hipMemcpyDeviceToHost, s)) != hipSuccess) {
    rc = fail(ISMC_EDEVICE, "hipMemcpyAsync", e);
    goto out;
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) rc = fail(ISMC_EDEVICE, "hipStreamSynchronize", e);
out:
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
