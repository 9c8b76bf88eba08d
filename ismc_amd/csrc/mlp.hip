// Surrogate MLP for the inverse solver: x (5 features) -> [256 SiLU] x 4 -> 4
// logit-transformed tallies, trained on this Monte Carlo's own output.
//
// All contractions run on the gfx950 matrix cores: BF16 operands, FP32
// accumulation (v_mfma_f32_16x16x32_bf16).  Master weights and Adam state
// are FP32; a BF16 copy of the weights feeds the MFMA.  There is no
// counterpart in the reference: it seeds the minimization solver.
//
// GEMM: C[M,N] = op(A)[M,K] * op(B)[K,N], 64x64 block tile, K step 32,
// four waves per block each owning a 32x32 quarter as 2x2 16x16 MFMA tiles.
// Tiles go through LDS stored K-contiguous (As[m][k], Bs[n][k]), so each
// lane's MFMA fragment (8 consecutive k of one row / column) is one 16-byte
// LDS read.  Transposed operands are handled in the global->LDS load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ismc.h"

namespace mlpk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int LDK = TK + 8;  // padded LDS row (bf16 elements)

__device__ __forceinline__ bf16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

enum Epi : int {
  EPI_STORE_F32 = 0,     // C (fp32)
  EPI_BIAS_SILU = 1,     // Z = C + b (fp32), A = silu(Z) (bf16)
  EPI_BIAS_F32 = 2,      // Y = C + b (fp32)
  EPI_SILU_GRAD = 3,     // dZ = C * silu'(Zprev) (bf16)
};

struct GemmArgs {
  int M, N, K;
  const bf16_t* A;
  int lda;
  const bf16_t* B;
  int ldb;
  float* C;        // fp32 out (EPI_STORE_F32 / Z / Y), row-major ldc
  int ldc;
  bf16_t* Cb;      // bf16 out (A or dZ), row-major ldc
  const float* bias;
  const float* Zp; // pre-activation for EPI_SILU_GRAD, row-major ldc
  int k_split;     // number of K slices (gridDim.z); C has k_split slabs of M x ldc
};

// op(A)[m][k] = A[m*lda + k] (TA = false) or A[k*lda + m] (TA = true)
// op(B)[k][n] = B[k*ldb + n] (TB = false) or B[n*ldb + k] (TB = true)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16_t As[TM][LDK];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[TN][LDK];
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kslice = g.K / g.k_split;
  const int kbeg = blockIdx.z * kslice, kend = kbeg + kslice;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // ---- A tile: 64 x 32 ----
    if (!TA) {
      const int r = t >> 2, c = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.A + (size_t)(m0 + r) * g.lda + k0 + c);
      *reinterpret_cast<uint4*>(&As[r][c]) = v;
    } else {
      const int kk = t >> 3, c = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.A + (size_t)(k0 + kk) * g.lda + m0 + c);
      const bf16_t* e = reinterpret_cast<const bf16_t*>(&v);
      for (int i = 0; i < 8; ++i) As[c + i][kk] = e[i];
    }
    // ---- B tile: 32 x 64, stored Bs[n][k] ----
    if (!TB) {
      const int kk = t >> 3, c = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.B + (size_t)(k0 + kk) * g.ldb + n0 + c);
      const bf16_t* e = reinterpret_cast<const bf16_t*>(&v);
      for (int i = 0; i < 8; ++i) Bs[c + i][kk] = e[i];
    } else {
      const int r = t >> 2, c = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.B + (size_t)(n0 + r) * g.ldb + k0 + c);
      *reinterpret_cast<uint4*>(&Bs[r][c]) = v;
    }
    __syncthreads();
    // ---- 2x2 MFMA 16x16x32 per wave ----
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i) {
      const uint4 va = *reinterpret_cast<const uint4*>(&As[wm + i * 16 + fr][fk]);
      a[i] = __builtin_bit_cast(bf16x8, va);
      const uint4 vb = *reinterpret_cast<const uint4*>(&Bs[wn + i * 16 + fr][fk]);
      b[i] = __builtin_bit_cast(bf16x8, vb);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // ---- epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each 16x16 tile ----
  float* C = g.C ? g.C + (size_t)blockIdx.z * g.M * g.ldc : nullptr;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        const float v = acc[i][j][r];
        const size_t o = (size_t)m * g.ldc + n;
        if (EPI == EPI_STORE_F32) {
          C[o] = v;
        } else if (EPI == EPI_BIAS_SILU) {
          const float z = v + g.bias[n];
          C[o] = z;
          g.Cb[o] = f2bf(z * sigmoid(z));
        } else if (EPI == EPI_BIAS_F32) {
          C[o] = v + g.bias[n];
        } else {
          const float z = g.Zp[o];
          const float s = sigmoid(z);
          g.Cb[o] = f2bf(v * s * (1.0f + z * (1.0f - s)));
        }
      }
}

// sum of k_split slabs (fixed order) -> out; also bias gradient rows
__global__ void reduce_slabs(const float* slabs, int k_split, size_t n, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < k_split; ++k) s += slabs[(size_t)k * n + i];
  out[i] = s;
}

// Partial bias gradients: partial[slab][n] = sum of dZ[b][n] (bf16) over the
// rows of one slab.  A block covers 64 columns (one row of 64 bf16 is a
// 128-byte coalesced load) with 4 row phases; the phases are combined in a
// fixed order, and reduce_slabs sums the slabs in a fixed order after it,
// so the result is deterministic.
constexpr int kBiasSlabs = 16;
__global__ void bias_partial(const bf16_t* dZ, int B, int N, int rows_per_slab, float* partial) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  const int r0 = blockIdx.y * rows_per_slab;
  const int r1 = min(B, r0 + rows_per_slab);
  float s = 0.f;
  if (n < N)
    for (int b = r0 + q; b < r1; b += 4) s += bf2f(dZ[(size_t)b * N + n]);
  red[q][c] = s;
  __syncthreads();
  if (q == 0 && n < N) partial[(size_t)blockIdx.y * N + n] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

// MSE on the first n_out columns: dY = 2 (Y - T) / (B n_out), loss partials per block
__global__ void mse_grad(const float* Y, int ldy, const float* T, int n_out, int B, bf16_t* dY,
                         float* loss_partial) {
  __shared__ float red[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < B) {
    for (int j = 0; j < ldy; ++j) {
      float d = 0.f;
      if (j < n_out) {
        d = Y[(size_t)b * ldy + j] - T[(size_t)b * n_out + j];
        l += d * d;
      }
      dY[(size_t)b * ldy + j] = f2bf(2.0f * d / (float)(B * n_out));
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// gather rows of the resident dataset into the batch buffers (features as
// padded bf16, targets fp32)
__global__ void gather_batch(const float* X, int n_in, const float* T, int n_out, const int* idx,
                             int B, int ldx, bf16_t* Xb, float* Tb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int r = idx[b];
  for (int j = 0; j < ldx; ++j) Xb[(size_t)b * ldx + j] = f2bf(j < n_in ? X[(size_t)r * n_in + j] : 0.f);
  for (int j = 0; j < n_out; ++j) Tb[(size_t)b * n_out + j] = T[(size_t)r * n_out + j];
}

__global__ void f32_to_bf16(const float* a, bf16_t* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// Adam (Kingma & Ba) with bias correction folded into the step size; keeps
// the BF16 copy of the weights in sync.  mask[i] == 0 marks padding that
// stays zero.
__global__ void adam_step(float* w, bf16_t* wb, const float* g, float* m, float* v,
                          const unsigned char* mask, size_t n, float lr_t, float b1, float b2,
                          float eps) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !mask[i]) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr_t * mi / (sqrtf(vi) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

}  // namespace mlpk

// ===========================================================================
// host side
// ===========================================================================
using namespace mlpk;

struct MlpLayer {
  int K, N;          // padded in/out
  int k_log, n_log;  // logical in/out
  size_t w_off, b_off;  // offsets into the flat fp32 parameter buffer
  size_t wb_off;        // offset of the bf16 weight copy
};

struct ismc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string err;
  int n_in = 0, n_out = 0, hidden = 0, depth = 0, max_batch = 0;
  int in_pad = 64, out_pad = 64;   // padded to the 64-wide GEMM tile
  std::vector<MlpLayer> L;
  size_t n_flat = 0, n_wb = 0;
  int64_t n_logical = 0;
  int k_split = 16;
  // device buffers
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  unsigned char* mask = nullptr;
  bf16_t* wb = nullptr;
  bf16_t* X0 = nullptr;                 // [maxB][in_pad]
  std::vector<float*> Z;                // per hidden layer [maxB][hidden]
  std::vector<bf16_t*> A;               // per hidden layer [maxB][hidden]
  float* Y = nullptr;                   // [maxB][out_pad]
  bf16_t* dA = nullptr;                 // scratch dZ buffers
  bf16_t* dB = nullptr;
  float* slabs = nullptr;               // split-K partials
  float* Tb = nullptr;                  // batch targets
  float* lossp = nullptr;
  float* lossd = nullptr;               // per-step loss partials of a training call
  size_t lossd_cap = 0;
  int* idx = nullptr;
  float* stage_x = nullptr;             // [maxB][n_in] upload staging
  float* stage_y = nullptr;             // [maxB][n_out]
  float* ds_x = nullptr;                // resident dataset
  float* ds_y = nullptr;
  int64_t ds_n = 0;
  int64_t steps = 0;
  float last_ms = 0.f;
};

static int mlp_err(ismc_mlp* p, int code, const char* what, hipError_t e = hipSuccess) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s%s%s", what, e != hipSuccess ? ": " : "",
           e != hipSuccess ? hipGetErrorString(e) : "");
  if (p) p->err = buf;
  return code;
}

#define MCHECK(p, call)                                                  \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) return mlp_err(p, ISMC_EDEVICE, #call, e_);     \
  } while (0)

template <typename T>
static int dalloc(ismc_mlp* p, T** ptr, size_t count) {
  hipError_t e = hipMalloc((void**)ptr, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return mlp_err(p, ISMC_ENOMEM, "hipMalloc", e);
  }
  return 0;
}

template <bool TA, bool TB, int EPI>
static int gemm(ismc_mlp* p, int M, int N, int K, const bf16_t* A, int lda, const bf16_t* B,
                int ldb, float* C, int ldc, bf16_t* Cb, const float* bias, const float* Zp,
                int k_split) {
  if (M % TM || N % TN || K % (TK * k_split)) return mlp_err(p, ISMC_EINVAL, "gemm shape not tileable");
  GemmArgs g{M, N, K, A, lda, B, ldb, C, ldc, Cb, bias, Zp, k_split};
  dim3 grid(N / TN, M / TM, k_split);
  gemm_kernel<TA, TB, EPI><<<grid, 256, 0, p->stream>>>(g);
  MCHECK(p, hipGetLastError());
  return 0;
}

static void mlp_free(ismc_mlp* p) {
  void* bufs[] = {p->w, p->g, p->m, p->v, p->mask, p->wb, p->X0, p->Y, p->dA, p->dB,
                  p->slabs, p->Tb, p->lossp, p->lossd, p->idx, p->ds_x, p->ds_y, p->stage_x, p->stage_y};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto z : p->Z)
    if (z) (void)hipFree(z);
  for (auto a : p->A)
    if (a) (void)hipFree(a);
}

// forward on the X0 batch buffer (B rows); leaves Z/A/Y populated
static int forward(ismc_mlp* p, int B) {
  int rc;
  const bf16_t* in = p->X0;
  int ld_in = p->in_pad;
  for (int l = 0; l < p->depth; ++l) {
    const MlpLayer& Ly = p->L[l];
    if ((rc = gemm<false, false, EPI_BIAS_SILU>(p, B, Ly.N, Ly.K, in, ld_in, p->wb + Ly.wb_off, Ly.N,
                                                  p->Z[l], Ly.N, p->A[l], p->w + Ly.b_off, nullptr, 1)))
      return rc;
    in = p->A[l];
    ld_in = Ly.N;
  }
  const MlpLayer& Lo = p->L[p->depth];
  return gemm<false, false, EPI_BIAS_F32>(p, B, Lo.N, Lo.K, in, ld_in, p->wb + Lo.wb_off, Lo.N, p->Y,
                                          Lo.N, nullptr, p->w + Lo.b_off, nullptr, 1);
}

// loss + backward on the current batch (targets in Tb); gradients into g
static int backward(ismc_mlp* p, int B, float* loss_host, float* loss_dev = nullptr) {
  int rc;
  const MlpLayer& Lo = p->L[p->depth];
  const int nb = (B + 255) / 256;
  mse_grad<<<nb, 256, 0, p->stream>>>(p->Y, Lo.N, p->Tb, p->n_out, B, p->dA, loss_dev ? loss_dev : p->lossp);
  MCHECK(p, hipGetLastError());
  bf16_t* dcur = p->dA;
  bf16_t* dnext = p->dB;
  for (int l = p->depth; l >= 0; --l) {
    const MlpLayer& Ly = p->L[l];
    const bf16_t* act = l == 0 ? p->X0 : p->A[l - 1];
    // dW = act^T dZ, split over the batch, partial slabs reduced in fixed order
    int ks = p->k_split;
    while (ks > 1 && B % (TK * ks)) ks >>= 1;
    if ((rc = gemm<true, false, EPI_STORE_F32>(p, Ly.K, Ly.N, B, act, Ly.K, dcur, Ly.N, p->slabs, Ly.N,
                                                 nullptr, nullptr, nullptr, ks)))
      return rc;
    const size_t nw = (size_t)Ly.K * Ly.N;
    reduce_slabs<<<(unsigned)((nw + 255) / 256), 256, 0, p->stream>>>(p->slabs, ks, nw, p->g + Ly.w_off);
    // bias gradient: column sums over kBiasSlabs row slabs (the split-K slabs
    // above are consumed by now and hold the partials: kBiasSlabs x N floats
    // fit in k_split x K x N), then a fixed-order sum
    {
      const int rows = (B + kBiasSlabs - 1) / kBiasSlabs;
      const int nslab = (B + rows - 1) / rows;
      bias_partial<<<dim3((unsigned)((Ly.N + 63) / 64), (unsigned)nslab), 256, 0, p->stream>>>(
          dcur, B, Ly.N, rows, p->slabs);
      reduce_slabs<<<(unsigned)((Ly.N + 255) / 256), 256, 0, p->stream>>>(p->slabs, nslab, (size_t)Ly.N,
                                                                           p->g + Ly.b_off);
    }
    MCHECK(p, hipGetLastError());
    if (l > 0) {
      // dZ_prev = (dZ W^T) * silu'(Z_prev)
      if ((rc = gemm<false, true, EPI_SILU_GRAD>(p, B, Ly.K, Ly.N, dcur, Ly.N, p->wb + Ly.wb_off, Ly.N,
                                                  nullptr, Ly.K, dnext, nullptr, p->Z[l - 1], 1)))
        return rc;
      std::swap(dcur, dnext);
    }
  }
  if (loss_host) {
    std::vector<float> lp(nb);
    MCHECK(p, hipMemcpyAsync(lp.data(), p->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, p->stream));
    MCHECK(p, hipStreamSynchronize(p->stream));
    double s = 0;
    for (float x : lp) s += x;
    *loss_host = (float)(s / ((double)B * p->n_out));
  }
  return 0;
}

static int upload_batch(ismc_mlp* p, const float* x, const float* y, int B) {
  // stage through persistent device buffers, rows 0..B-1 in order
  std::vector<int> id(B);
  for (int i = 0; i < B; ++i) id[i] = i;
  float* tx = p->stage_x;
  float* ty = y ? p->stage_y : nullptr;
  hipError_t e = hipMemcpyAsync(tx, x, sizeof(float) * B * p->n_in, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess && y)
    e = hipMemcpyAsync(ty, y, sizeof(float) * B * p->n_out, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->idx, id.data(), sizeof(int) * B, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) {
    gather_batch<<<(B + 255) / 256, 256, 0, p->stream>>>(tx, p->n_in, ty ? ty : tx, y ? p->n_out : 0,
                                                         p->idx, B, p->in_pad, p->X0, p->Tb);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return mlp_err(p, ISMC_EDEVICE, "batch upload", e);
  return 0;
}

// one Adam step on every layer from the gradients in p->g
static int adam_all(ismc_mlp* p, float lr, float b1, float b2, float eps) {
  p->steps++;
  const float bc1 = 1.f - powf(b1, (float)p->steps), bc2 = 1.f - powf(b2, (float)p->steps);
  const float lr_t = lr * sqrtf(bc2) / bc1;
  for (const MlpLayer& Ly : p->L) {
    const size_t nw = (size_t)Ly.K * Ly.N;
    adam_step<<<(unsigned)((nw + 255) / 256), 256, 0, p->stream>>>(
        p->w + Ly.w_off, p->wb + Ly.wb_off, p->g + Ly.w_off, p->m + Ly.w_off, p->v + Ly.w_off,
        p->mask + Ly.w_off, nw, lr_t, b1, b2, eps);
    adam_step<<<(unsigned)((Ly.N + 255) / 256), 256, 0, p->stream>>>(
        p->w + Ly.b_off, nullptr, p->g + Ly.b_off, p->m + Ly.b_off, p->v + Ly.b_off, p->mask + Ly.b_off,
        (size_t)Ly.N, lr_t, b1, b2, eps);
  }
  MCHECK(p, hipGetLastError());
  return 0;
}

static int check_batch(ismc_mlp* p, int B) {
  if (B <= 0 || B > p->max_batch || B % TM)
    return mlp_err(p, ISMC_EINVAL, "batch must be a positive multiple of 64 and <= max_batch");
  return 0;
}

extern "C" {

int ismc_mlp_create(int device, int32_t n_in, int32_t hidden, int32_t depth, int32_t n_out,
                    int32_t max_batch, ismc_mlp** out) {
  if (!out) return ISMC_EINVAL;
  *out = nullptr;
  if (n_in < 1 || n_in > 32 || n_out < 1 || n_out > 64 || hidden < 64 || hidden % 64 || depth < 1 ||
      depth > 16 || max_batch < 64 || max_batch % 64)
    return ISMC_EINVAL;
  ismc_mlp* p = new ismc_mlp();
  p->device = device;
  p->n_in = n_in;
  p->n_out = n_out;
  p->hidden = hidden;
  p->depth = depth;
  p->max_batch = max_batch;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&p->e0) != hipSuccess || hipEventCreate(&p->e1) != hipSuccess) {
    delete p;
    return ISMC_EDEVICE;
  }
  size_t off = 0, woff = 0;
  for (int l = 0; l <= depth; ++l) {
    MlpLayer Ly;
    Ly.K = l == 0 ? p->in_pad : hidden;
    Ly.N = l == depth ? p->out_pad : hidden;
    Ly.k_log = l == 0 ? n_in : hidden;
    Ly.n_log = l == depth ? n_out : hidden;
    Ly.w_off = off;
    off += (size_t)Ly.K * Ly.N;
    Ly.b_off = off;
    off += Ly.N;
    Ly.wb_off = woff;
    woff += (size_t)Ly.K * Ly.N;
    p->n_logical += (int64_t)Ly.k_log * Ly.n_log + Ly.n_log;
    p->L.push_back(Ly);
  }
  p->n_flat = off;
  p->n_wb = woff;
  int rc = 0;
  const size_t B = max_batch;
  rc = rc ? rc : dalloc(p, &p->w, off);
  rc = rc ? rc : dalloc(p, &p->g, off);
  rc = rc ? rc : dalloc(p, &p->m, off);
  rc = rc ? rc : dalloc(p, &p->v, off);
  rc = rc ? rc : dalloc(p, &p->mask, off);
  rc = rc ? rc : dalloc(p, &p->wb, woff);
  rc = rc ? rc : dalloc(p, &p->X0, B * p->in_pad);
  rc = rc ? rc : dalloc(p, &p->Y, B * p->out_pad);
  rc = rc ? rc : dalloc(p, &p->dA, B * std::max(hidden, p->out_pad));
  rc = rc ? rc : dalloc(p, &p->dB, B * std::max(hidden, p->out_pad));
  rc = rc ? rc : dalloc(p, &p->slabs, (size_t)p->k_split * hidden * std::max(hidden, p->out_pad));
  rc = rc ? rc : dalloc(p, &p->Tb, B * n_out);
  rc = rc ? rc : dalloc(p, &p->lossp, (B + 255) / 256);
  rc = rc ? rc : dalloc(p, &p->idx, B);
  rc = rc ? rc : dalloc(p, &p->stage_x, B * n_in);
  rc = rc ? rc : dalloc(p, &p->stage_y, B * n_out);
  p->Z.assign(depth, nullptr);
  p->A.assign(depth, nullptr);
  for (int l = 0; l < depth && !rc; ++l) {
    rc = dalloc(p, &p->Z[l], B * hidden);
    rc = rc ? rc : dalloc(p, &p->A[l], B * hidden);
  }
  if (rc) {
    mlp_free(p);
    delete p;
    return rc;
  }
  // zero everything; the mask marks logical (non-padding) parameters
  std::vector<unsigned char> mask(off, 0);
  for (const MlpLayer& Ly : p->L) {
    for (int k = 0; k < Ly.k_log; ++k)
      for (int n = 0; n < Ly.n_log; ++n) mask[Ly.w_off + (size_t)k * Ly.N + n] = 1;
    for (int n = 0; n < Ly.n_log; ++n) mask[Ly.b_off + n] = 1;
  }
  hipError_t e = hipMemcpy(p->mask, mask.data(), off, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->w, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(p->g, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(p->m, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(p->v, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(p->wb, 0, woff * 2);
  if (e != hipSuccess) {
    mlp_free(p);
    delete p;
    return ISMC_EDEVICE;
  }
  *out = p;
  return 0;
}

void ismc_mlp_destroy(ismc_mlp* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  (void)hipStreamSynchronize(p->stream);
  mlp_free(p);
  (void)hipEventDestroy(p->e0);
  (void)hipEventDestroy(p->e1);
  (void)hipStreamDestroy(p->stream);
  delete p;
}

const char* ismc_mlp_last_error(const ismc_mlp* p) { return p ? p->err.c_str() : "null mlp"; }

int64_t ismc_mlp_n_params(const ismc_mlp* p) { return p ? p->n_logical : -1; }

// logical layout: for each layer, W[in][out] row-major then b[out]
int ismc_mlp_set_params(ismc_mlp* p, const float* flat) {
  if (!p || !flat) return ISMC_EINVAL;
  MCHECK(p, hipSetDevice(p->device));
  std::vector<float> h(p->n_flat, 0.f);
  size_t o = 0;
  for (const MlpLayer& Ly : p->L) {
    for (int k = 0; k < Ly.k_log; ++k)
      for (int n = 0; n < Ly.n_log; ++n) h[Ly.w_off + (size_t)k * Ly.N + n] = flat[o++];
    for (int n = 0; n < Ly.n_log; ++n) h[Ly.b_off + n] = flat[o++];
  }
  MCHECK(p, hipMemcpyAsync(p->w, h.data(), p->n_flat * 4, hipMemcpyHostToDevice, p->stream));
  for (const MlpLayer& Ly : p->L) {
    const size_t n = (size_t)Ly.K * Ly.N;
    f32_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, p->stream>>>(p->w + Ly.w_off, p->wb + Ly.wb_off, n);
  }
  MCHECK(p, hipGetLastError());
  MCHECK(p, hipMemsetAsync(p->m, 0, p->n_flat * 4, p->stream));
  MCHECK(p, hipMemsetAsync(p->v, 0, p->n_flat * 4, p->stream));
  MCHECK(p, hipStreamSynchronize(p->stream));
  p->steps = 0;
  return 0;
}

static int get_flat(ismc_mlp* p, const float* dev, float* flat) {
  std::vector<float> h(p->n_flat);
  MCHECK(p, hipMemcpyAsync(h.data(), dev, p->n_flat * 4, hipMemcpyDeviceToHost, p->stream));
  MCHECK(p, hipStreamSynchronize(p->stream));
  size_t o = 0;
  for (const MlpLayer& Ly : p->L) {
    for (int k = 0; k < Ly.k_log; ++k)
      for (int n = 0; n < Ly.n_log; ++n) flat[o++] = h[Ly.w_off + (size_t)k * Ly.N + n];
    for (int n = 0; n < Ly.n_log; ++n) flat[o++] = h[Ly.b_off + n];
  }
  return 0;
}

int ismc_mlp_get_params(ismc_mlp* p, float* flat) {
  if (!p || !flat) return ISMC_EINVAL;
  MCHECK(p, hipSetDevice(p->device));
  return get_flat(p, p->w, flat);
}

int ismc_mlp_forward(ismc_mlp* p, const float* x, int32_t n, float* y) {
  if (!p || !x || !y || n < 0) return ISMC_EINVAL;
  MCHECK(p, hipSetDevice(p->device));
  int rc;
  // host buffer for one batch's padded outputs, sized by this call's rows
  const int32_t rows = std::min<int32_t>(p->max_batch, ((n + TM - 1) / TM) * TM);
  std::vector<float> yb((size_t)rows * p->out_pad);
  for (int32_t s = 0; s < n; s += p->max_batch) {
    const int cnt = std::min<int32_t>(p->max_batch, n - s);
    const int B = ((cnt + TM - 1) / TM) * TM;
    std::vector<float> xb((size_t)B * p->n_in, 0.f);
    memcpy(xb.data(), x + (size_t)s * p->n_in, sizeof(float) * cnt * p->n_in);
    if ((rc = upload_batch(p, xb.data(), nullptr, B))) return rc;
    MCHECK(p, hipEventRecord(p->e0, p->stream));
    if ((rc = forward(p, B))) return rc;
    MCHECK(p, hipEventRecord(p->e1, p->stream));
    MCHECK(p, hipMemcpyAsync(yb.data(), p->Y, sizeof(float) * B * p->out_pad, hipMemcpyDeviceToHost, p->stream));
    MCHECK(p, hipStreamSynchronize(p->stream));
    MCHECK(p, hipEventElapsedTime(&p->last_ms, p->e0, p->e1));
    for (int i = 0; i < cnt; ++i)
      for (int j = 0; j < p->n_out; ++j) y[(size_t)(s + i) * p->n_out + j] = yb[(size_t)i * p->out_pad + j];
  }
  return 0;
}

int ismc_mlp_gradients(ismc_mlp* p, const float* x, const float* y, int32_t B, float* grads,
                       float* loss) {
  if (!p || !x || !y || !grads) return ISMC_EINVAL;
  int rc;
  if ((rc = check_batch(p, B))) return rc;
  MCHECK(p, hipSetDevice(p->device));
  if ((rc = upload_batch(p, x, y, B))) return rc;
  if ((rc = forward(p, B))) return rc;
  if ((rc = backward(p, B, loss))) return rc;
  return get_flat(p, p->g, grads);
}

int ismc_mlp_load_dataset(ismc_mlp* p, const float* x, const float* y, int64_t n) {
  if (!p || !x || !y || n < p->max_batch || n > INT32_MAX)
    return mlp_err(p, ISMC_EINVAL, "dataset needs max_batch .. 2^31-1 rows and non-null arrays");
  MCHECK(p, hipSetDevice(p->device));
  // the new buffers first: if one allocation fails the context keeps its
  // previous dataset and index buffer intact
  float* nx = nullptr;
  float* ny = nullptr;
  int* nidx = nullptr;
  int rc = dalloc(p, &nx, (size_t)n * p->n_in);
  rc = rc ? rc : dalloc(p, &ny, (size_t)n * p->n_out);
  rc = rc ? rc : dalloc(p, &nidx, (size_t)n);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpy(nx, x, sizeof(float) * n * p->n_in, hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(ny, y, sizeof(float) * n * p->n_out, hipMemcpyHostToDevice);
  if (rc || e != hipSuccess) {
    if (nx) (void)hipFree(nx);
    if (ny) (void)hipFree(ny);
    if (nidx) (void)hipFree(nidx);
    return rc ? rc : mlp_err(p, ISMC_EDEVICE, "dataset upload", e);
  }
  if (p->ds_x) (void)hipFree(p->ds_x);
  if (p->ds_y) (void)hipFree(p->ds_y);
  if (p->idx) (void)hipFree(p->idx);
  p->ds_x = nx;
  p->ds_y = ny;
  p->idx = nidx;
  p->ds_n = n;
  return 0;
}

// `steps_per_epoch` batches of `B` rows drawn in the order `perm` (host,
// steps_per_epoch * B indices per epoch, epochs consecutive); returns the
// mean loss of each epoch.  Adam: lr, beta1, beta2, eps.
int ismc_mlp_train(ismc_mlp* p, const int32_t* perm, int32_t epochs, int32_t steps_per_epoch,
                   int32_t B, float lr, float b1, float b2, float eps, float* epoch_loss,
                   float* total_ms) {
  if (!p || !perm || epochs < 1 || steps_per_epoch < 1 || !p->ds_x) return ISMC_EINVAL;
  int rc;
  if ((rc = check_batch(p, B))) return rc;
  if ((int64_t)steps_per_epoch * B > p->ds_n) return mlp_err(p, ISMC_EINVAL, "epoch larger than the dataset");
  // every index must name a resident row: gather_batch reads ds_x / ds_y at
  // them, so one outside 0 .. ds_n-1 would read beyond the dataset
  const size_t n_idx = (size_t)epochs * steps_per_epoch * B;
  for (size_t i = 0; i < n_idx; ++i)
    if (perm[i] < 0 || (int64_t)perm[i] >= p->ds_n)
      return mlp_err(p, ISMC_EINVAL, "perm holds an index outside the resident dataset");
  MCHECK(p, hipSetDevice(p->device));
  const int nb = (B + 255) / 256;
  std::vector<float> lp((size_t)nb * steps_per_epoch);
  if (p->lossd_cap < lp.size()) {
    if (p->lossd) (void)hipFree(p->lossd);
    p->lossd = nullptr;
    p->lossd_cap = 0;
    if ((rc = dalloc(p, &p->lossd, lp.size()))) return rc;
    p->lossd_cap = lp.size();
  }
  float* lossd = p->lossd;
  float ms_total = 0.f;
  for (int e = 0; e < epochs; ++e) {
    MCHECK(p, hipMemcpyAsync(p->idx, perm + (size_t)e * steps_per_epoch * B,
                             sizeof(int32_t) * (size_t)steps_per_epoch * B, hipMemcpyHostToDevice, p->stream));
    MCHECK(p, hipEventRecord(p->e0, p->stream));
    double esum = 0;
    for (int s = 0; s < steps_per_epoch; ++s) {
      gather_batch<<<nb, 256, 0, p->stream>>>(p->ds_x, p->n_in, p->ds_y, p->n_out, p->idx + (size_t)s * B, B,
                                              p->in_pad, p->X0, p->Tb);
      MCHECK(p, hipGetLastError());
      if ((rc = forward(p, B)) || (rc = backward(p, B, nullptr, lossd + (size_t)s * nb))) return rc;
      if ((rc = adam_all(p, lr, b1, b2, eps))) return rc;
    }
    MCHECK(p, hipEventRecord(p->e1, p->stream));
    MCHECK(p, hipMemcpyAsync(lp.data(), lossd, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, p->stream));
    MCHECK(p, hipStreamSynchronize(p->stream));
    for (float x : lp) esum += x / ((double)B * p->n_out);
    float ms = 0.f;
    MCHECK(p, hipEventElapsedTime(&ms, p->e0, p->e1));
    ms_total += ms;
    if (epoch_loss) epoch_loss[e] = (float)(esum / steps_per_epoch);
  }
  if (total_ms) *total_ms = ms_total;
  return 0;
}

// One Adam step on a host batch x[B][n_in], y[B][n_out] (targets in the
// output's units): upload, forward, backward, update; *loss (if non-NULL)
// is the batch's mean squared error before the step.
int ismc_mlp_train_step(ismc_mlp* p, const float* x, const float* y, int32_t B, float lr, float b1,
                        float b2, float eps, float* loss) {
  if (!p || !x || !y) return ISMC_EINVAL;
  int rc;
  if ((rc = check_batch(p, B))) return rc;
  MCHECK(p, hipSetDevice(p->device));
  if ((rc = upload_batch(p, x, y, B))) return rc;
  if ((rc = forward(p, B)) || (rc = backward(p, B, loss))) return rc;
  if ((rc = adam_all(p, lr, b1, b2, eps))) return rc;
  MCHECK(p, hipStreamSynchronize(p->stream));
  return 0;
}

int ismc_mlp_last_ms(ismc_mlp* p, float* ms) {
  if (!p || !ms) return ISMC_EINVAL;
  *ms = p->last_ms;
  return 0;
}

}  // extern "C"
