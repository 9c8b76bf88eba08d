// Surrogate MLP for the inverse solver: x (5 features) -> [hidden SiLU] x depth
// -> 4 logit-transformed tallies, trained on this Monte Carlo's own output.
// There is no counterpart in the reference: it seeds the minimization solver.
//
// Every contraction runs on the gfx950 matrix cores: v_mfma_f32_16x16x32_bf16,
// BF16 operands, FP32 accumulation.  Master weights and Adam state are FP32;
// two BF16 copies of every weight matrix feed the MFMA, W^T [out][in] for
// the forward and W [in][out] for the backward, so that each lane's B
// fragment (8 consecutive k of one column) is one 16-byte load in both.
//
// A training step is four launches:
//   1. mlp_forward   one block per 16 batch rows runs all layers: the rows'
//                    activations stay in LDS (A fragments), the weights come
//                    from L2 as B fragments; each hidden layer's epilogue
//                    adds the bias and applies SiLU, keeps the pre-activation
//                    (fragment order) and writes the activation transposed
//                    ([unit][row], the weight-gradient kernel's A operand);
//                    the output layer's epilogue forms the MSE gradient and
//                    the block's squared-error sum.
//   2. mlp_backward  the same row blocks run the gradient back through the
//                    layers (dZ W^T, times SiLU'), writing every layer's dZ
//                    transposed ([unit][row], the weight-gradient kernel's B
//                    operand).
//   3. mlp_dweight   every layer's weight gradient A^T dZ at once: one wave
//                    per 32 x 32 tile of one layer and one slice of the
//                    batch rows; the waves of each layer's first tile row
//                    also sum their dZ fragments (the bias gradient); one
//                    slab of partial sums per row slice.
//   4. mlp_adam      each parameter sums its slabs in a fixed order and takes
//                    one Adam step, refreshing both BF16 copies.
// Every sum has a fixed order, so training is deterministic.  Inputs and
// outputs are padded to 32 (one MFMA K step each), with the padding masked
// out of Adam.
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

constexpr int kRows = 16;       // batch rows per block of the fused kernels
constexpr int kInPad = 32;      // input features padded to one MFMA K step
constexpr int kOutPad = 32;     // outputs padded to one MFMA K step (the backward's first)
constexpr int kMaxLayers = 17;  // at most 16 hidden layers and the output layer
constexpr int kMaxSlices = 8;   // batch-row slices of the weight-gradient kernel

__device__ __forceinline__ bf16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ uint2 pack4(const uint16_t h[4]) {
  uint2 pk;
  pk.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
  pk.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
  return pk;
}

// Device view of one network's parameters.  Layer l maps K[l] inputs to
// N[l] outputs; in the FP32 flat buffer its W [K][N] (row-major) starts at
// w_off[l] and its bias at b_off[l]; both BF16 copies start at wb_off[l].
struct Net {
  int n_in, n_out, H, depth;
  const float* w;
  const bf16_t* wt;  // W^T [N][K] per layer
  const bf16_t* wb;  // W   [K][N] per layer
  int K[kMaxLayers], N[kMaxLayers];
  int w_off[kMaxLayers], b_off[kMaxLayers], wb_off[kMaxLayers];
};

// One wave's strip of TILES 16 x 16 output tiles (columns n0 + 16 t) over
// KSTEPS steps of 32: A fragments from the block's rows in LDS (row-major,
// k contiguous), B fragments from global memory, Bt[n][k] (k contiguous,
// leading dimension ldb).  Every B fragment of the strip is requested
// before the first MFMA, so that the strip waits for one L2 latency rather
// than one per K step (a block has one wave per SIMD or two: no other
// wave hides the latency).  MFMA operand layout (16x16x32): lane l holds
// row / column l % 16, k = 8 (l / 16) .. + 7; the result holds column
// l % 16, rows 4 (l / 16) .. + 3.
template <int TILES, int KSTEPS, int LDA>
__device__ __forceinline__ void strip(const bf16_t (*A)[LDA], const bf16_t* __restrict__ Bt, int ldb,
                                      int n0, int lane, f32x4 acc[TILES]) {
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  bf16x8 bv[KSTEPS][TILES];
#pragma unroll
  for (int s = 0; s < KSTEPS; ++s)
#pragma unroll
    for (int t = 0; t < TILES; ++t) bv[s][t] = load8(Bt + (size_t)(n0 + t * 16 + fr) * ldb + s * 32 + fk);
#pragma unroll
  for (int t = 0; t < TILES; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < KSTEPS; ++s) {
    const bf16x8 av = load8(&A[fr][s * 32 + fk]);
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[s][t], acc[t], 0, 0, 0);
  }
}

// waves per block of the fused kernels: two 16-column tiles per wave of a
// hidden layer, at most 8 waves (two per SIMD)
template <int H> struct Waves { static constexpr int value = H / 32 < 8 ? H / 32 : 8; };

struct FwdArgs {
  Net net;
  int B;               // rows of the batch (a multiple of kRows)
  int n_valid;         // rows that hold data (the rest get zero inputs)
  const float* X;      // source rows [*][n_in]
  const float* T;      // targets [*][n_out] (training) or null
  const int* idx;      // row -> source row (null: the row itself)
  float* Y;            // inference: outputs [n_valid][n_out], or null
  // training outputs, null for inference
  float* Zf;           // pre-activations of the hidden layers, fragment order
  bf16_t* XT;          // inputs transposed [kInPad][B]
  bf16_t* AT;          // hidden activations transposed [depth][H][B]
  bf16_t* dYT;         // output gradient transposed [kOutPad][B]
  float* loss_part;    // [B / kRows] squared-error sums
  float grad_scale;    // 2 / (B n_out)
};

// pre-activation fragment of (layer l, row block rb, column tile c): 64 lanes x 4
__device__ __forceinline__ size_t zf_index(int l, int rb, int c, int nrb, int H, int lane) {
  return ((((size_t)l * nrb + rb) * (H / 16) + c) * 64 + lane) * 4;
}

template <int H>
__global__ __launch_bounds__(64 * Waves<H>::value) void mlp_forward(FwdArgs a) {
  constexpr int LDA = H + 8;
  constexpr int NW = Waves<H>::value;
  constexpr int TW = H / (16 * NW);  // 16-column tiles per wave of a hidden layer
  __shared__ __attribute__((aligned(16))) bf16_t act[2][kRows][LDA];
  __shared__ float red[2];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int rb = blockIdx.x, row0 = rb * kRows, nrb = gridDim.x;
  const int fr = lane & 15, cr = (lane >> 4) * 4;
  const Net& net = a.net;
  for (int e = tid; e < kRows * kInPad; e += 64 * NW) {
    const int r = e / kInPad, k = e % kInPad;
    const int row = row0 + r;
    float v = 0.f;
    if (k < net.n_in && row < a.n_valid) v = a.X[(size_t)(a.idx ? a.idx[row] : row) * net.n_in + k];
    const bf16_t h = f2bf(v);
    act[0][r][k] = h;
    if (a.XT) a.XT[(size_t)k * a.B + row] = h;
  }
  __syncthreads();
  int cur = 0;
  for (int l = 0; l < net.depth; ++l) {
    const int n0 = wave * (H / NW);
    const bf16_t* Wt = net.wt + net.wb_off[l];
    const float* bias = net.w + net.b_off[l];
    f32x4 acc[TW];
    if (l == 0)
      strip<TW, kInPad / 32, LDA>(act[cur], Wt, kInPad, n0, lane, acc);
    else
      strip<TW, H / 32, LDA>(act[cur], Wt, H, n0, lane, acc);
#pragma unroll
    for (int t = 0; t < TW; ++t) {
      const int n = n0 + t * 16 + fr;
      const float b = bias[n];
      f32x4 z;
      uint16_t h[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        z[r] = acc[t][r] + b;
        h[r] = f2bf(z[r] * sigmoid(z[r]));
        act[cur ^ 1][cr + r][n] = h[r];
      }
      if (a.Zf) {
        *reinterpret_cast<f32x4*>(a.Zf + zf_index(l, rb, n0 / 16 + t, nrb, H, lane)) = z;
        *reinterpret_cast<uint2*>(a.AT + ((size_t)l * H + n) * a.B + row0 + cr) = pack4(h);
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  // output layer: 32 padded columns, one tile each for waves 0 and 1
  if (wave < kOutPad / 16) {
    const int L = net.depth;
    const int n = wave * 16 + fr;
    f32x4 acc[1];
    strip<1, H / 32, LDA>(act[cur], net.wt + net.wb_off[L], H, wave * 16, lane, acc);
    const float b = net.w[net.b_off[L] + n];
    uint16_t h[4];
    float sq = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + cr + r;
      const float y = acc[0][r] + b;
      const bool live = n < net.n_out && row < a.n_valid;
      if (a.Y && live) a.Y[(size_t)row * net.n_out + n] = y;
      float d = 0.f;
      if (a.T && live) d = y - a.T[(size_t)(a.idx ? a.idx[row] : row) * net.n_out + n];
      sq += d * d;
      h[r] = f2bf(d * a.grad_scale);
    }
    if (a.dYT) *reinterpret_cast<uint2*>(a.dYT + (size_t)n * a.B + row0 + cr) = pack4(h);
    // the wave's sum in a fixed order (butterfly), then wave 0 + wave 1
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if (lane == 0) red[wave] = sq;
  }
  if (a.loss_part) {
    __syncthreads();
    if (tid == 0) a.loss_part[rb] = red[0] + red[1];
  }
}

struct BwdArgs {
  Net net;
  int B;
  const float* Zf;
  const bf16_t* dYT;
  bf16_t* dZT;  // [depth][H][B]
};

template <int H>
__global__ __launch_bounds__(64 * Waves<H>::value) void mlp_backward(BwdArgs a) {
  constexpr int LDA = H + 8;
  constexpr int NW = Waves<H>::value;
  constexpr int TW = H / (16 * NW);
  __shared__ __attribute__((aligned(16))) bf16_t g[2][kRows][LDA];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int rb = blockIdx.x, row0 = rb * kRows, nrb = gridDim.x;
  const int fr = lane & 15, cr = (lane >> 4) * 4;
  const Net& net = a.net;
  for (int e = tid; e < kRows * kOutPad; e += 64 * NW) {
    const int r = e % kRows, k = e / kRows;
    g[0][r][k] = a.dYT[(size_t)k * a.B + row0 + r];
  }
  __syncthreads();
  int cur = 0;
  for (int l = net.depth; l >= 1; --l) {
    // gradient at layer l-1's output: dZ_l [rows][N_l] W_l^T, with the B
    // fragments read from W_l [K_l][N_l] (row k_in holds one lane's N_l values)
    const int n0 = wave * (H / NW);
    const bf16_t* W = net.wb + net.wb_off[l];
    f32x4 acc[TW];
    if (l == net.depth)
      strip<TW, kOutPad / 32, LDA>(g[cur], W, kOutPad, n0, lane, acc);
    else
      strip<TW, H / 32, LDA>(g[cur], W, H, n0, lane, acc);
#pragma unroll
    for (int t = 0; t < TW; ++t) {
      const int n = n0 + t * 16 + fr;
      const f32x4 z = *reinterpret_cast<const f32x4*>(a.Zf + zf_index(l - 1, rb, n0 / 16 + t, nrb, H, lane));
      uint16_t h[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float s = sigmoid(z[r]);
        h[r] = f2bf(acc[t][r] * s * (1.0f + z[r] * (1.0f - s)));
        g[cur ^ 1][cr + r][n] = h[r];
      }
      *reinterpret_cast<uint2*>(a.dZT + ((size_t)(l - 1) * H + n) * a.B + row0 + cr) = pack4(h);
    }
    __syncthreads();
    cur ^= 1;
  }
}

struct DwArgs {
  Net net;
  int B;
  int slices;
  int rows_per_slice;
  int n_flat;
  int tile0[kMaxLayers + 1];  // first 32 x 32 tile of each layer (prefix sums)
  const bf16_t* XT;
  const bf16_t* AT;
  const bf16_t* dZT;
  const bf16_t* dYT;
  float* slab;  // [slices][n_flat]
};

// Q K steps of one wave's 32 x 32 weight-gradient tile: every fragment is
// requested before the first MFMA; with brow, the dZ fragments' column sums
// too (the bias gradient)
template <int Q>
__device__ __forceinline__ void dw_steps(const bf16_t* __restrict__ A, const bf16_t* __restrict__ G, int B,
                                         int m0, int n0, int k0, int fr, int fk, bool brow, f32x4 acc[2][2],
                                         float bsum[2]) {
  bf16x8 av[Q][2], bv[Q][2];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      av[q][i] = load8(A + (size_t)(m0 + i * 16 + fr) * B + k0 + q * 32 + fk);
      bv[q][i] = load8(G + (size_t)(n0 + i * 16 + fr) * B + k0 + q * 32 + fk);
    }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[q][i], bv[q][j], acc[i][j], 0, 0, 0);
    if (brow) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        float w = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) w += (float)bv[q][j][e];
        bsum[j] += w;
      }
    }
  }
}

// Weight gradients dW_l = A_{l-1}^T dZ_l of every layer: one wave per
// 32 x 32 tile of one layer and one slice of the batch rows (rows of A^T
// and dZ^T are k-contiguous, so both fragments are 16-byte loads).  The
// waves of a layer's first tile row also sum their dZ fragments: the bias
// gradient's partial over the slice.  Every word of a slice's slab (the
// padding included) is written by exactly one wave.
__global__ __launch_bounds__(256) void mlp_dweight(DwArgs a) {
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int L = a.net.depth + 1;
  const int tiles = a.tile0[L];
  const int s = gw / tiles;
  if (s >= a.slices) return;
  int t = gw - s * tiles;
  int l = 0;
  while (t >= a.tile0[l + 1]) ++l;
  t -= a.tile0[l];
  const int N = a.net.N[l];
  const int m0 = (t / (N / 32)) * 32, n0 = (t % (N / 32)) * 32;
  const bf16_t* A = l == 0 ? a.XT : a.AT + (size_t)(l - 1) * a.net.H * a.B;
  const bf16_t* G = l == a.net.depth ? a.dYT : a.dZT + (size_t)l * a.net.H * a.B;
  const int fr = lane & 15, fk = (lane >> 4) * 8, cr = (lane >> 4) * 4;
  const int r0 = s * a.rows_per_slice;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[2] = {0.f, 0.f};
  const bool brow = (m0 == 0);
  // Q K steps (32 Q rows) of fragments in flight at a time: four while a
  // slice has 128 rows left, then one (a slice is a multiple of 32 rows)
  const int r1 = r0 + a.rows_per_slice;
  int k0 = r0;
  for (; k0 + 128 <= r1; k0 += 128) dw_steps<4>(A, G, a.B, m0, n0, k0, fr, fk, brow, acc, bsum);
  for (; k0 < r1; k0 += 32) dw_steps<1>(A, G, a.B, m0, n0, k0, fr, fk, brow, acc, bsum);
  float* out = a.slab + (size_t)s * a.n_flat;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[a.net.w_off[l] + (size_t)(m0 + i * 16 + cr + r) * N + n0 + j * 16 + fr] = acc[i][j][r];
  if (brow) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      // lanes fr, fr + 16, fr + 32, fr + 48 hold parts of the same column
      float q = bsum[j];
      q += __shfl_xor(q, 16);
      q += __shfl_xor(q, 32);
      if (lane < 16) out[a.net.b_off[l] + n0 + j * 16 + fr] = q;
    }
  }
}

struct AdamArgs {
  Net net;
  float* w;
  float* g;               // gradient output (when !adam)
  float* m;
  float* v;
  const unsigned char* mask;
  const float* slab;
  int slices;
  int n_flat;
  int adam;
  float lr_t, b1, b2, eps;
  bf16_t* wt;
  bf16_t* wb;
};

// refresh both BF16 copies of flat parameter i if it is a weight
__device__ __forceinline__ void put_bf16(const Net& net, int i, float wi, bf16_t* wt, bf16_t* wb) {
  for (int l = 0; l <= net.depth; ++l) {
    const int rel = i - net.w_off[l];
    if (rel >= 0 && rel < net.K[l] * net.N[l]) {
      const int k = rel / net.N[l], n = rel - k * net.N[l];
      const bf16_t h = f2bf(wi);
      wb[net.wb_off[l] + rel] = h;
      wt[net.wb_off[l] + n * net.K[l] + k] = h;
      return;
    }
  }
}

// Slab sums in a fixed order, then Adam (Kingma & Ba) with the bias
// correction folded into lr_t; or, with adam == 0, the gradient itself.
__global__ void mlp_adam(AdamArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_flat) return;
  float gi = 0.f;
  for (int s = 0; s < a.slices; ++s) gi += a.slab[(size_t)s * a.n_flat + i];
  if (!a.adam) {
    a.g[i] = gi;
    return;
  }
  if (!a.mask[i]) return;
  const float mi = a.b1 * a.m[i] + (1.f - a.b1) * gi;
  const float vi = a.b2 * a.v[i] + (1.f - a.b2) * gi * gi;
  a.m[i] = mi;
  a.v[i] = vi;
  const float wi = a.w[i] - a.lr_t * mi / (sqrtf(vi) + a.eps);
  a.w[i] = wi;
  put_bf16(a.net, i, wi, a.wt, a.wb);
}

// both BF16 copies from the FP32 weights (after set_params)
__global__ void mlp_sync_bf16(Net net, const float* w, int n_flat, bf16_t* wt, bf16_t* wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_flat) put_bf16(net, i, w[i], wt, wb);
}

__global__ void fill_zero(float* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

}  // namespace mlpk

// ===========================================================================
// host side
// ===========================================================================
using namespace mlpk;

struct ismc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string err;
  int n_in = 0, n_out = 0, hidden = 0, depth = 0, max_batch = 0;
  int n_flat = 0, n_wb = 0;
  int64_t n_logical = 0;
  Net net{};
  // parameters
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  unsigned char* mask = nullptr;
  bf16_t *wt = nullptr, *wb = nullptr;
  // host-batch staging and inference output (max_batch rows)
  float* stage_x = nullptr;   // [max_batch][n_in]
  float* stage_y = nullptr;   // [max_batch][n_out]
  float* Y = nullptr;         // [max_batch][n_out]
  float* lossp = nullptr;     // [max_batch / kRows]
  // training buffers, allocated on first use for a batch size
  int train_B = 0;
  float* Zf = nullptr;
  bf16_t *XT = nullptr, *AT = nullptr, *dZT = nullptr, *dYT = nullptr;
  float* slab = nullptr;
  float* lossd = nullptr;     // per-step loss partials of a training call
  size_t lossd_cap = 0;
  int* idx = nullptr;         // one epoch's dataset order
  size_t idx_cap = 0;
  float* ds_x = nullptr;      // resident dataset
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

static void free_train(ismc_mlp* p) {
  void* bufs[] = {p->Zf, p->XT, p->AT, p->dZT, p->dYT, p->slab};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  p->Zf = nullptr;
  p->XT = p->AT = p->dZT = p->dYT = nullptr;
  p->slab = nullptr;
  p->train_B = 0;
}

static void mlp_free(ismc_mlp* p) {
  free_train(p);
  void* bufs[] = {p->w, p->g, p->m, p->v, p->mask, p->wt, p->wb, p->stage_x, p->stage_y, p->Y,
                  p->lossp, p->lossd, p->idx, p->ds_x, p->ds_y};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
}

// training buffers for batches of B rows
static int ensure_train(ismc_mlp* p, int B) {
  if (p->train_B >= B) return 0;
  free_train(p);
  const size_t H = p->hidden, D = p->depth;
  int rc = dalloc(p, &p->Zf, D * H * B);
  rc = rc ? rc : dalloc(p, &p->XT, (size_t)kInPad * B);
  rc = rc ? rc : dalloc(p, &p->AT, D * H * B);
  rc = rc ? rc : dalloc(p, &p->dZT, D * H * B);
  rc = rc ? rc : dalloc(p, &p->dYT, (size_t)kOutPad * B);
  rc = rc ? rc : dalloc(p, &p->slab, (size_t)kMaxSlices * p->n_flat);
  if (rc) {
    free_train(p);
    return rc;
  }
  p->train_B = B;
  return 0;
}

template <int H>
static void launch_fwd(ismc_mlp* p, const FwdArgs& a) {
  mlp_forward<H><<<a.B / kRows, 64 * Waves<H>::value, 0, p->stream>>>(a);
}
template <int H>
static void launch_bwd(ismc_mlp* p, const BwdArgs& a) {
  mlp_backward<H><<<a.B / kRows, 64 * Waves<H>::value, 0, p->stream>>>(a);
}

static int run_forward(ismc_mlp* p, FwdArgs a) {
  a.net = p->net;
  switch (p->hidden) {
    case 64: launch_fwd<64>(p, a); break;
    case 128: launch_fwd<128>(p, a); break;
    case 256: launch_fwd<256>(p, a); break;
    default: launch_fwd<512>(p, a); break;
  }
  MCHECK(p, hipGetLastError());
  return 0;
}

// forward with every training output (rows of X through idx, or in order;
// targets T), backward, and the weight gradients into the slabs; returns
// the number of slabs, or a negative status
static int forward_backward(ismc_mlp* p, int B, const float* X, const float* T, const int* idx,
                            float* loss_part) {
  int rc;
  if ((rc = ensure_train(p, B))) return rc;
  FwdArgs f{};
  f.B = B;
  f.n_valid = B;
  f.X = X;
  f.T = T;
  f.idx = idx;
  f.Zf = p->Zf;
  f.XT = p->XT;
  f.AT = p->AT;
  f.dYT = p->dYT;
  f.loss_part = loss_part;
  f.grad_scale = 2.0f / (float)(B * p->n_out);
  if ((rc = run_forward(p, f))) return rc;
  BwdArgs b{};
  b.net = p->net;
  b.B = B;
  b.Zf = p->Zf;
  b.dYT = p->dYT;
  b.dZT = p->dZT;
  switch (p->hidden) {
    case 64: launch_bwd<64>(p, b); break;
    case 128: launch_bwd<128>(p, b); break;
    case 256: launch_bwd<256>(p, b); break;
    default: launch_bwd<512>(p, b); break;
  }
  MCHECK(p, hipGetLastError());
  DwArgs d{};
  d.net = p->net;
  d.B = B;
  int slices = kMaxSlices;
  while (slices > 1 && B % (32 * slices)) slices >>= 1;
  d.slices = slices;
  d.rows_per_slice = B / slices;
  d.n_flat = p->n_flat;
  d.tile0[0] = 0;
  for (int l = 0; l <= p->depth; ++l) d.tile0[l + 1] = d.tile0[l] + (p->net.K[l] / 32) * (p->net.N[l] / 32);
  d.XT = p->XT;
  d.AT = p->AT;
  d.dZT = p->dZT;
  d.dYT = p->dYT;
  d.slab = p->slab;
  const int waves = d.tile0[p->depth + 1] * slices;
  mlp_dweight<<<(waves + 3) / 4, 256, 0, p->stream>>>(d);
  MCHECK(p, hipGetLastError());
  return slices;
}

// the slabs' sums into p->g (adam == false), or one Adam step
static int reduce_step(ismc_mlp* p, int slices, bool adam, float lr, float b1, float b2, float eps) {
  AdamArgs a{};
  a.net = p->net;
  a.w = p->w;
  a.g = p->g;
  a.m = p->m;
  a.v = p->v;
  a.mask = p->mask;
  a.slab = p->slab;
  a.slices = slices;
  a.n_flat = p->n_flat;
  a.adam = adam ? 1 : 0;
  if (adam) {
    p->steps++;
    const float bc1 = 1.f - powf(b1, (float)p->steps), bc2 = 1.f - powf(b2, (float)p->steps);
    a.lr_t = lr * sqrtf(bc2) / bc1;
  }
  a.b1 = b1;
  a.b2 = b2;
  a.eps = eps;
  a.wt = p->wt;
  a.wb = p->wb;
  mlp_adam<<<(p->n_flat + 255) / 256, 256, 0, p->stream>>>(a);
  MCHECK(p, hipGetLastError());
  return 0;
}

static int check_batch(ismc_mlp* p, int B) {
  if (B <= 0 || B > p->max_batch || B % 64)
    return mlp_err(p, ISMC_EINVAL, "batch must be a positive multiple of 64 and <= max_batch");
  return 0;
}

// host rows x[rows][n_in] (and y[rows][n_out]) into the staging buffers
static int upload(ismc_mlp* p, const float* x, const float* y, int rows) {
  MCHECK(p, hipMemcpyAsync(p->stage_x, x, sizeof(float) * rows * p->n_in, hipMemcpyHostToDevice, p->stream));
  if (y)
    MCHECK(p, hipMemcpyAsync(p->stage_y, y, sizeof(float) * rows * p->n_out, hipMemcpyHostToDevice, p->stream));
  return 0;
}

// mean squared error of the batch from the per-block sums in p->lossp
static int loss_of(ismc_mlp* p, int B, float* loss) {
  const int nb = B / kRows;
  std::vector<float> lp(nb);
  MCHECK(p, hipMemcpyAsync(lp.data(), p->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, p->stream));
  MCHECK(p, hipStreamSynchronize(p->stream));
  double s = 0;
  for (float x : lp) s += x;
  *loss = (float)(s / ((double)B * p->n_out));
  return 0;
}

extern "C" {

int ismc_mlp_create(int device, int32_t n_in, int32_t hidden, int32_t depth, int32_t n_out,
                    int32_t max_batch, ismc_mlp** out) {
  if (!out) return ISMC_EINVAL;
  *out = nullptr;
  if (n_in < 1 || n_in > kInPad || n_out < 1 || n_out > kOutPad ||
      !(hidden == 64 || hidden == 128 || hidden == 256 || hidden == 512) || depth < 1 ||
      depth > kMaxLayers - 1 || max_batch < 64 || max_batch % 64)
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
    ismc_mlp_destroy(p);
    return ISMC_EDEVICE;
  }
  Net& net = p->net;
  net.n_in = n_in;
  net.n_out = n_out;
  net.H = hidden;
  net.depth = depth;
  int off = 0, woff = 0;
  for (int l = 0; l <= depth; ++l) {
    net.K[l] = l == 0 ? kInPad : hidden;
    net.N[l] = l == depth ? kOutPad : hidden;
    const int k_log = l == 0 ? n_in : hidden, n_log = l == depth ? n_out : hidden;
    net.w_off[l] = off;
    off += net.K[l] * net.N[l];
    net.b_off[l] = off;
    off += net.N[l];
    net.wb_off[l] = woff;
    woff += net.K[l] * net.N[l];
    p->n_logical += (int64_t)k_log * n_log + n_log;
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
  rc = rc ? rc : dalloc(p, &p->wt, woff);
  rc = rc ? rc : dalloc(p, &p->wb, woff);
  rc = rc ? rc : dalloc(p, &p->stage_x, B * n_in);
  rc = rc ? rc : dalloc(p, &p->stage_y, B * n_out);
  rc = rc ? rc : dalloc(p, &p->Y, B * n_out);
  rc = rc ? rc : dalloc(p, &p->lossp, B / kRows);
  if (rc) {
    ismc_mlp_destroy(p);
    return rc;
  }
  net.w = p->w;
  net.wt = p->wt;
  net.wb = p->wb;
  // zero everything; the mask marks logical (non-padding) parameters
  std::vector<unsigned char> mask(off, 0);
  for (int l = 0; l <= depth; ++l) {
    const int k_log = l == 0 ? n_in : hidden, n_log = l == depth ? n_out : hidden;
    for (int k = 0; k < k_log; ++k)
      for (int n = 0; n < n_log; ++n) mask[net.w_off[l] + (size_t)k * net.N[l] + n] = 1;
    for (int n = 0; n < n_log; ++n) mask[net.b_off[l] + n] = 1;
  }
  hipError_t e = hipMemcpy(p->mask, mask.data(), off, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->w, 0, (size_t)off * 4);
  if (e == hipSuccess) e = hipMemset(p->g, 0, (size_t)off * 4);
  if (e == hipSuccess) e = hipMemset(p->m, 0, (size_t)off * 4);
  if (e == hipSuccess) e = hipMemset(p->v, 0, (size_t)off * 4);
  if (e == hipSuccess) e = hipMemset(p->wt, 0, (size_t)woff * 2);
  if (e == hipSuccess) e = hipMemset(p->wb, 0, (size_t)woff * 2);
  if (e != hipSuccess) {
    ismc_mlp_destroy(p);
    return ISMC_EDEVICE;
  }
  *out = p;
  return 0;
}

void ismc_mlp_destroy(ismc_mlp* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  mlp_free(p);
  if (p->e0) (void)hipEventDestroy(p->e0);
  if (p->e1) (void)hipEventDestroy(p->e1);
  if (p->stream) (void)hipStreamDestroy(p->stream);
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
  const Net& net = p->net;
  for (int l = 0; l <= p->depth; ++l) {
    const int k_log = l == 0 ? p->n_in : p->hidden, n_log = l == p->depth ? p->n_out : p->hidden;
    for (int k = 0; k < k_log; ++k)
      for (int n = 0; n < n_log; ++n) h[net.w_off[l] + (size_t)k * net.N[l] + n] = flat[o++];
    for (int n = 0; n < n_log; ++n) h[net.b_off[l] + n] = flat[o++];
  }
  MCHECK(p, hipMemcpyAsync(p->w, h.data(), (size_t)p->n_flat * 4, hipMemcpyHostToDevice, p->stream));
  const unsigned nb = (unsigned)((p->n_flat + 255) / 256);
  mlp_sync_bf16<<<nb, 256, 0, p->stream>>>(p->net, p->w, p->n_flat, p->wt, p->wb);
  fill_zero<<<nb, 256, 0, p->stream>>>(p->m, (size_t)p->n_flat);
  fill_zero<<<nb, 256, 0, p->stream>>>(p->v, (size_t)p->n_flat);
  MCHECK(p, hipGetLastError());
  MCHECK(p, hipStreamSynchronize(p->stream));
  p->steps = 0;
  return 0;
}

static int get_flat(ismc_mlp* p, const float* dev, float* flat) {
  std::vector<float> h(p->n_flat);
  MCHECK(p, hipMemcpyAsync(h.data(), dev, (size_t)p->n_flat * 4, hipMemcpyDeviceToHost, p->stream));
  MCHECK(p, hipStreamSynchronize(p->stream));
  size_t o = 0;
  const Net& net = p->net;
  for (int l = 0; l <= p->depth; ++l) {
    const int k_log = l == 0 ? p->n_in : p->hidden, n_log = l == p->depth ? p->n_out : p->hidden;
    for (int k = 0; k < k_log; ++k)
      for (int n = 0; n < n_log; ++n) flat[o++] = h[net.w_off[l] + (size_t)k * net.N[l] + n];
    for (int n = 0; n < n_log; ++n) flat[o++] = h[net.b_off[l] + n];
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
  for (int32_t s = 0; s < n; s += p->max_batch) {
    const int cnt = std::min<int32_t>(p->max_batch, n - s);
    const int B = ((cnt + kRows - 1) / kRows) * kRows;
    if ((rc = upload(p, x + (size_t)s * p->n_in, nullptr, cnt))) return rc;
    FwdArgs f{};
    f.B = B;
    f.n_valid = cnt;
    f.X = p->stage_x;
    f.Y = p->Y;
    MCHECK(p, hipEventRecord(p->e0, p->stream));
    if ((rc = run_forward(p, f))) return rc;
    MCHECK(p, hipEventRecord(p->e1, p->stream));
    MCHECK(p, hipMemcpyAsync(y + (size_t)s * p->n_out, p->Y, sizeof(float) * cnt * p->n_out,
                             hipMemcpyDeviceToHost, p->stream));
    MCHECK(p, hipStreamSynchronize(p->stream));
    MCHECK(p, hipEventElapsedTime(&p->last_ms, p->e0, p->e1));
  }
  return 0;
}

int ismc_mlp_gradients(ismc_mlp* p, const float* x, const float* y, int32_t B, float* grads,
                       float* loss) {
  if (!p || !x || !y || !grads) return ISMC_EINVAL;
  int rc;
  if ((rc = check_batch(p, B))) return rc;
  MCHECK(p, hipSetDevice(p->device));
  if ((rc = upload(p, x, y, B))) return rc;
  const int slices = forward_backward(p, B, p->stage_x, p->stage_y, nullptr, p->lossp);
  if (slices < 0) return slices;
  if ((rc = reduce_step(p, slices, false, 0.f, 0.f, 0.f, 0.f))) return rc;
  if (loss && (rc = loss_of(p, B, loss))) return rc;
  return get_flat(p, p->g, grads);
}

int ismc_mlp_load_dataset(ismc_mlp* p, const float* x, const float* y, int64_t n) {
  if (!p || !x || !y || n < p->max_batch || n > INT32_MAX)
    return mlp_err(p, ISMC_EINVAL, "dataset needs max_batch .. 2^31-1 rows and non-null arrays");
  MCHECK(p, hipSetDevice(p->device));
  // the new buffers first: if one allocation fails the context keeps its
  // previous dataset intact
  float* nx = nullptr;
  float* ny = nullptr;
  int rc = dalloc(p, &nx, (size_t)n * p->n_in);
  rc = rc ? rc : dalloc(p, &ny, (size_t)n * p->n_out);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpy(nx, x, sizeof(float) * n * p->n_in, hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(ny, y, sizeof(float) * n * p->n_out, hipMemcpyHostToDevice);
  if (rc || e != hipSuccess) {
    if (nx) (void)hipFree(nx);
    if (ny) (void)hipFree(ny);
    return rc ? rc : mlp_err(p, ISMC_EDEVICE, "dataset upload", e);
  }
  if (p->ds_x) (void)hipFree(p->ds_x);
  if (p->ds_y) (void)hipFree(p->ds_y);
  p->ds_x = nx;
  p->ds_y = ny;
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
  // every index must name a resident row: the forward reads ds_x / ds_y at
  // them, so one outside 0 .. ds_n-1 would read beyond the dataset
  const size_t per_epoch = (size_t)steps_per_epoch * B;
  const size_t n_idx = (size_t)epochs * per_epoch;
  for (size_t i = 0; i < n_idx; ++i)
    if (perm[i] < 0 || (int64_t)perm[i] >= p->ds_n)
      return mlp_err(p, ISMC_EINVAL, "perm holds an index outside the resident dataset");
  MCHECK(p, hipSetDevice(p->device));
  const int nb = B / kRows;
  std::vector<float> lp((size_t)nb * steps_per_epoch);
  if (p->lossd_cap < lp.size()) {
    if (p->lossd) (void)hipFree(p->lossd);
    p->lossd = nullptr;
    p->lossd_cap = 0;
    if ((rc = dalloc(p, &p->lossd, lp.size()))) return rc;
    p->lossd_cap = lp.size();
  }
  if (p->idx_cap < per_epoch) {
    if (p->idx) (void)hipFree(p->idx);
    p->idx = nullptr;
    p->idx_cap = 0;
    if ((rc = dalloc(p, &p->idx, per_epoch))) return rc;
    p->idx_cap = per_epoch;
  }
  if ((rc = ensure_train(p, B))) return rc;
  float ms_total = 0.f;
  for (int e = 0; e < epochs; ++e) {
    MCHECK(p, hipMemcpyAsync(p->idx, perm + (size_t)e * per_epoch, sizeof(int32_t) * per_epoch,
                             hipMemcpyHostToDevice, p->stream));
    MCHECK(p, hipEventRecord(p->e0, p->stream));
    for (int s = 0; s < steps_per_epoch; ++s) {
      const int slices = forward_backward(p, B, p->ds_x, p->ds_y, p->idx + (size_t)s * B,
                                          p->lossd + (size_t)s * nb);
      if (slices < 0) return slices;
      if ((rc = reduce_step(p, slices, true, lr, b1, b2, eps))) return rc;
    }
    MCHECK(p, hipEventRecord(p->e1, p->stream));
    MCHECK(p, hipMemcpyAsync(lp.data(), p->lossd, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, p->stream));
    MCHECK(p, hipStreamSynchronize(p->stream));
    double esum = 0;
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
  if ((rc = upload(p, x, y, B))) return rc;
  const int slices = forward_backward(p, B, p->stage_x, p->stage_y, nullptr, p->lossp);
  if (slices < 0) return slices;
  if ((rc = reduce_step(p, slices, true, lr, b1, b2, eps))) return rc;
  if (loss && (rc = loss_of(p, B, loss))) return rc;
  MCHECK(p, hipStreamSynchronize(p->stream));
  return 0;
}

int ismc_mlp_last_ms(ismc_mlp* p, float* ms) {
  if (!p || !ms) return ISMC_EINVAL;
  *ms = p->last_ms;
  return 0;
}

}  // extern "C"
