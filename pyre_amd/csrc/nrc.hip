// Neural radiance cache on gfx950 BF16 MFMA (no counterpart in the reference).
//
// Network: 13 raw features per query (position, outgoing direction, normal,
// roughness, albedo) -> 128-wide encoding (12 sin/cos frequencies per position
// axis on scene-normalised coordinates, real SH l<=3 of direction and of the
// normal, roughness, albedo, a constant 1 that acts as the first layer's bias,
// zero padding) -> 5 hidden layers of 64 with ReLU -> 16 outputs, RGB in
// 0..2. No biases. BF16 weights and activations, FP32 accumulation, FP32
// master weights with Adam. Loss: relative L2 (pred - t)^2 / (sg(pred)^2 + 0.01).
//
// All GEMMs use v_mfma_f32_16x16x32_bf16 in the transposed form
// Z^T[out x batch] = W^T[out x in] . X^T[in x batch]: the A operand is a
// weight fragment, the B operand 8 consecutive features of one query, so
// activations stay sample-major ([query][feature]) everywhere. Operand lane
// layout (lane l, g = l / 16, r = l % 16): A[r][8g + j], B[8g + j][r],
// C[4g + i][r] for j < 8, i < 4.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "nrc.hpp"

using namespace pyre;

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kLayers = 6;
constexpr int kIn[kLayers] = {128, 64, 64, 64, 64, 64};
constexpr int kOut[kLayers] = {64, 64, 64, 64, 64, 16};

__host__ __device__ constexpr int param_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += kIn[i] * kOut[i];
  return o;
}
constexpr int kParams = param_offset(kLayers);  // 25600
// fragment buffers: per layer, (out/16) x ceil(in/32) fragments of 64 lanes x 8 bf16
__host__ __device__ constexpr int fwd_frags(int l) { return (kOut[l] / 16) * (kIn[l] / 32); }
__host__ __device__ constexpr int bwd_frags(int l) { return (kIn[l] / 16) * ((kOut[l] + 31) / 32); }
__host__ __device__ constexpr int fwd_frag_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += fwd_frags(i);
  return o;
}
__host__ __device__ constexpr int bwd_frag_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += bwd_frags(i);
  return o;
}
constexpr int kFwdFrags = fwd_frag_offset(kLayers);  // 50
constexpr int kBwdFrags = bwd_frag_offset(kLayers);

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }  // round to nearest even
__device__ __forceinline__ float from_bf16(__bf16 x) { return (float)x; }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------ encoding
__device__ __forceinline__ float sh4(int i, float x, float y, float z) {
  float xx = x * x, yy = y * y, zz = z * z;
  switch (i) {
    case 0: return 0.28209479177387814f;
    case 1: return -0.48860251190291987f * y;
    case 2: return 0.48860251190291987f * z;
    case 3: return -0.48860251190291987f * x;
    case 4: return 1.0925484305920792f * x * y;
    case 5: return -1.0925484305920792f * y * z;
    case 6: return 0.94617469575755997f * zz - 0.31539156525251999f;
    case 7: return -1.0925484305920792f * x * z;
    case 8: return 0.54627421529603959f * (xx - yy);
    case 9: return 0.59004358992664352f * y * (-3.0f * xx + yy);
    case 10: return 2.8906114426405538f * x * y * z;
    case 11: return 0.45704579946446572f * y * (1.0f - 5.0f * zz);
    case 12: return 0.3731763325901154f * z * (5.0f * zz - 3.0f);
    case 13: return 0.45704579946446572f * x * (1.0f - 5.0f * zz);
    case 14: return 1.4453057213202769f * z * (xx - yy);
    default: return 0.59004358992664352f * x * (-xx + 3.0f * yy);
  }
}

struct Query {
  float px, py, pz;  // normalised position
  float dx, dy, dz, nx, ny, nz;
  float rough, ar, ag, ab;
};

__device__ __forceinline__ Query load_query(const float* f, const NrcBounds& b) {
  Query q;
  q.px = (f[0] - b.lo[0]) * b.inv_ext[0];
  q.py = (f[1] - b.lo[1]) * b.inv_ext[1];
  q.pz = (f[2] - b.lo[2]) * b.inv_ext[2];
  q.dx = f[3];
  q.dy = f[4];
  q.dz = f[5];
  q.nx = f[6];
  q.ny = f[7];
  q.nz = f[8];
  q.rough = f[9];
  q.ar = f[10];
  q.ag = f[11];
  q.ab = f[12];
  return q;
}

// Feature f of the 128-wide encoding (selects instead of indexed arrays, so
// the query stays in registers when f differs across lanes).
__device__ __forceinline__ float encode_feature(int f, const Query& q) {
  if (f < 72) {
    const int a = f / 24, k = (f % 24) / 2;
    const float pa = a == 0 ? q.px : (a == 1 ? q.py : q.pz);
    const float arg = ldexpf(pa, k);  // 2^k * p, exact
    return (f & 1) ? cospif(arg) : sinpif(arg);
  }
  if (f < 88) return sh4(f - 72, q.dx, q.dy, q.dz);
  if (f < 104) return sh4(f - 88, q.nx, q.ny, q.nz);
  if (f == 104) return q.rough;
  if (f == 105) return q.ar;
  if (f == 106) return q.ag;
  if (f == 107) return q.ab;
  return f == 108 ? 1.0f : 0.0f;
}

// B fragment of the encoding for K block kb: features 32kb + 8g + j.
__device__ __forceinline__ bf16x8 encode_frag(const Query& q, int kb, int g) {
  bf16x8 v;
  for (int j = 0; j < 8; ++j) v[j] = to_bf16(encode_feature(32 * kb + 8 * g + j, q));
  return v;
}

__global__ void k_encode(const float* feats, uint32_t n, NrcBounds b, __bf16* x) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Query q = load_query(feats + 13 * (size_t)i, b);
  for (int f = 0; f < 128; ++f) x[128 * (size_t)i + f] = to_bf16(encode_feature(f, q));
}

// ------------------------------------------------- weights -> fragments
// Forward fragment (l, ot, kb), lane (g, r): W_l[32kb + 8g + j][16ot + r].
// Backward fragment (l, it, kb), lane (g, r): W_l[16it + r][32kb + 8g + j]
// (zero where the K index runs past the layer's outputs).
__global__ void k_make_frags(const float* w, bf16x8* fwd, bf16x8* bwd) {
  int frag = blockIdx.x, lane = threadIdx.x;
  int g = lane >> 4, r = lane & 15;
  if (frag < kFwdFrags) {
    int l = 0;
    while (l + 1 < kLayers && frag >= fwd_frag_offset(l + 1)) ++l;
    int k = frag - fwd_frag_offset(l);
    int nkb = kIn[l] / 32;
    int ot = k / nkb, kb = k % nkb;
    const float* W = w + param_offset(l);
    bf16x8 v;
    for (int j = 0; j < 8; ++j) v[j] = to_bf16(W[(32 * kb + 8 * g + j) * kOut[l] + 16 * ot + r]);
    fwd[(size_t)frag * 64 + lane] = v;
  } else {
    int bf = frag - kFwdFrags;
    int l = 0;
    while (l + 1 < kLayers && bf >= bwd_frag_offset(l + 1)) ++l;
    int k = bf - bwd_frag_offset(l);
    int nkb = (kOut[l] + 31) / 32;
    int it = k / nkb, kb = k % nkb;
    const float* W = w + param_offset(l);
    bf16x8 v;
    for (int j = 0; j < 8; ++j) {
      int o = 32 * kb + 8 * g + j;
      v[j] = to_bf16(o < kOut[l] ? W[(16 * it + r) * kOut[l] + o] : 0.0f);
    }
    bwd[(size_t)bf * 64 + lane] = v;
  }
}

// ---------------------------------------------------------- inference
// Fused forward of whole 16-query tiles: every layer's weight fragments sit
// in LDS (51 KB), hidden activations go through a 2 KB per-wave LDS tile.
constexpr int kInferWaves = 8;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kInferWaves) void k_infer(const bf16x8* __restrict__ wfrag,
                                                            const float* __restrict__ feats,
                                                            const uint32_t* __restrict__ rows, uint32_t n_cap,
                                                            const uint32_t* __restrict__ n_dev, NrcBounds b,
                                                            float* __restrict__ out_rgb) {
  __shared__ bf16x8 sw[kFwdFrags * 64];
  // query count: n_cap, or the device-side count (clamped to n_cap) when given
  const uint32_t n = n_dev ? min(*n_dev, n_cap) : n_cap;
  __shared__ __attribute__((aligned(16))) __bf16 sact[kInferWaves][16][64 + 8];
  for (int i = threadIdx.x; i < kFwdFrags * 64; i += blockDim.x) sw[i] = wfrag[i];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, r = lane & 15;
  const uint32_t ntiles = (n + 15) / 16;
  for (uint32_t tile = blockIdx.x * kInferWaves + wave; tile < ntiles; tile += gridDim.x * kInferWaves) {
    const uint32_t qi = tile * 16 + r;
    const bool valid = qi < n;
    const uint32_t row = valid ? (rows ? rows[qi] : qi) : (rows ? rows[0] : 0u);
    Query q = load_query(feats + 13 * (size_t)row, b);
    // layer 0: K = 128 from the encoding
    f32x4 acc[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) acc[ot] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      bf16x8 bx = encode_frag(q, kb, g);
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) acc[ot] = mfma(sw[(fwd_frag_offset(0) + ot * 4 + kb) * 64 + lane], bx, acc[ot]);
    }
#pragma unroll
    for (int l = 1; l < kLayers; ++l) {
      // ReLU(previous) -> LDS tile [query][feature]
#pragma unroll
      for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int i = 0; i < 4; ++i) sact[wave][r][16 * ot + 4 * g + i] = to_bf16(fmaxf(acc[ot][i], 0.0f));
      wave_lds_sync();
      bf16x8 bx[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) bx[kb] = *(const bf16x8*)&sact[wave][r][32 * kb + 8 * g];
      wave_lds_sync();
      const int nout = kOut[l] / 16;
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) {
        acc[ot] = f32x4{0, 0, 0, 0};
        if (ot < nout) {
#pragma unroll
          for (int kb = 0; kb < 2; ++kb)
            acc[ot] = mfma(sw[(fwd_frag_offset(l) + ot * 2 + kb) * 64 + lane], bx[kb], acc[ot]);
        }
      }
    }
    if (valid && g == 0) {
      out_rgb[3 * (size_t)qi] = acc[0][0];
      out_rgb[3 * (size_t)qi + 1] = acc[0][1];
      out_rgb[3 * (size_t)qi + 2] = acc[0][2];
    }
  }
}

// ---------------------------------------------------------- training
// Forward layer: Y[n][out] = act(X[n][in] W). One wave per 16-query tile.
template <bool RELU_BF16_OUT>
__global__ __launch_bounds__(256) void k_fwd_layer(const bf16x8* __restrict__ wfrag, int layer,
                                                   const __bf16* __restrict__ x, uint32_t n, __bf16* y_bf16,
                                                   float* y_f32) {
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile * 16 >= n) return;
  const int in = kIn[layer], out = kOut[layer];
  const int nkb = in / 32, nout = out / 16;
  const __bf16* xr = x + (size_t)(tile * 16 + r) * in;
  bf16x8 bx[4];
  for (int kb = 0; kb < nkb; ++kb) bx[kb] = *(const bf16x8*)(xr + 32 * kb + 8 * g);
  for (int ot = 0; ot < nout; ++ot) {
    f32x4 acc = {0, 0, 0, 0};
    for (int kb = 0; kb < nkb; ++kb) acc = mfma(wfrag[(size_t)(fwd_frag_offset(layer) + ot * nkb + kb) * 64 + lane], bx[kb], acc);
    size_t o = (size_t)(tile * 16 + r) * out + 16 * ot + 4 * g;
    if (RELU_BF16_OUT) {
      for (int i = 0; i < 4; ++i) y_bf16[o + i] = to_bf16(fmaxf(acc[i], 0.0f));
    } else {
      for (int i = 0; i < 4; ++i) y_f32[o + i] = acc[i];
    }
  }
}

// Relative L2 loss: dP (bf16, 16 wide, zero past RGB) and per-block loss sums.
__global__ __launch_bounds__(256) void k_loss(const float* pred, const float* target, uint32_t n, uint32_t n_real,
                                              __bf16* dp, float* loss_part) {
  __shared__ float red[256];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (i < n) {
    for (int c = 0; c < 16; ++c) {
      float gval = 0.0f;
      if (c < 3 && i < n_real) {
        float p = pred[16 * (size_t)i + c], t = target[3 * (size_t)i + c];
        float denom = p * p + 0.01f;
        float d = p - t;
        s += d * d / denom;
        gval = 2.0f * d / denom / (3.0f * (float)n_real);
      }
      dp[16 * (size_t)i + c] = to_bf16(gval);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ_{l-1} = (dZ_l W_l^T) * (X_l > 0): dX^T[in x n] = W_l[in x out] . dZ_l^T.
__global__ __launch_bounds__(256) void k_bwd_dx(const bf16x8* __restrict__ bfrag, int layer,
                                                const __bf16* __restrict__ dz, const __bf16* __restrict__ xl,
                                                uint32_t n, __bf16* dz_prev) {
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile * 16 >= n) return;
  const int in = kIn[layer], out = kOut[layer];
  const int nkb = (out + 31) / 32, nit = in / 16;
  const uint32_t q = tile * 16 + r;
  bf16x8 bz[2];
  for (int kb = 0; kb < nkb; ++kb) {
    if (32 * kb + 8 * g < out) {
      bz[kb] = *(const bf16x8*)(dz + (size_t)q * out + 32 * kb + 8 * g);
    } else {
      for (int j = 0; j < 8; ++j) bz[kb][j] = to_bf16(0.0f);
    }
  }
  for (int it = 0; it < nit; ++it) {
    f32x4 acc = {0, 0, 0, 0};
    for (int kb = 0; kb < nkb; ++kb) acc = mfma(bfrag[(size_t)(bwd_frag_offset(layer) + it * nkb + kb) * 64 + lane], bz[kb], acc);
    size_t o = (size_t)q * in + 16 * it + 4 * g;
    for (int i = 0; i < 4; ++i) dz_prev[o + i] = to_bf16(from_bf16(xl[o + i]) > 0.0f ? acc[i] : 0.0f);
  }
}

// dW_l = X_l^T dZ_l over a chunk of kDwChunk queries per block (partial sums,
// reduced deterministically by k_dw_reduce). A = X^T tile gathered from LDS,
// B = dZ tile gathered from LDS.
constexpr int kDwChunk = 128;
__global__ __launch_bounds__(256) void k_bwd_dw(int layer, const __bf16* __restrict__ xl,
                                                const __bf16* __restrict__ dz, uint32_t n, float* partial) {
  __shared__ __bf16 sx[kDwChunk][128 + 2];
  __shared__ __bf16 sz[kDwChunk][64 + 2];
  const int in = kIn[layer], out = kOut[layer];
  const uint32_t q0 = blockIdx.x * kDwChunk;
  for (int i = threadIdx.x; i < kDwChunk * in; i += blockDim.x) {
    int s = i / in, f = i % in;
    sx[s][f] = (q0 + s < n) ? xl[(size_t)(q0 + s) * in + f] : to_bf16(0.0f);
  }
  for (int i = threadIdx.x; i < kDwChunk * out; i += blockDim.x) {
    int s = i / out, f = i % out;
    sz[s][f] = (q0 + s < n) ? dz[(size_t)(q0 + s) * out + f] : to_bf16(0.0f);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const int nit = in / 16, not_ = out / 16;
  float* dst = partial + (size_t)blockIdx.x * in * out;
  for (int t = wave; t < nit * not_; t += 4) {
    int it = t / not_, ot = t % not_;
    f32x4 acc = {0, 0, 0, 0};
    for (int kb = 0; kb < kDwChunk / 32; ++kb) {
      bf16x8 a, bb;
      for (int j = 0; j < 8; ++j) {
        a[j] = sx[32 * kb + 8 * g + j][16 * it + r];   // A[in row][query k]
        bb[j] = sz[32 * kb + 8 * g + j][16 * ot + r];  // B[query k][out col]
      }
      acc = mfma(a, bb, acc);
    }
    for (int i = 0; i < 4; ++i) dst[(size_t)(16 * it + 4 * g + i) * out + 16 * ot + r] = acc[i];
  }
}

__global__ void k_dw_reduce(const float* partial, uint32_t nchunks, int size, float* grad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  float s = 0.0f;
  for (uint32_t c = 0; c < nchunks; ++c) s += partial[(size_t)c * size + i];
  grad[i] = s;
}

// Adam (Kingma & Ba) on the FP32 master weights; the RGB-only output columns
// and the padding rows of layer 0 have zero gradient and stay fixed.
__global__ void k_adam(float* w, const float* grad, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gr = grad[i];
  float mm = b1 * m[i] + (1.0f - b1) * gr;
  float vv = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mm;
  v[i] = vv;
  float mh = mm / c1, vh = vv / c2;
  w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct pyre_nrc {
  pyre_ctx* ctx = nullptr;
  pyre_nrc_params prm{};
  NrcBounds bounds{};
  float* w = nullptr;  // FP32 master weights, layer by layer, [in][out] row-major
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  bf16x8* wf = nullptr;  // forward fragments
  bf16x8* wb = nullptr;  // backward fragments
  uint32_t step = 0;
  float last_loss = 0.0f;
  // training scratch, sized for the largest batch seen
  uint32_t cap = 0;
  __bf16* acts = nullptr;  // layer inputs: 128 + 5 * 64 features per query
  __bf16* dz0 = nullptr;
  __bf16* dz1 = nullptr;
  float* pred = nullptr;
  float* part = nullptr;
  float* loss_part = nullptr;
  float* feats = nullptr;
  float* targets = nullptr;
};

namespace pyre {

int nrc_param_count() { return kParams; }

void nrc_refresh_frags(pyre_nrc* N, hipStream_t st) {
  hipLaunchKernelGGL(k_make_frags, dim3(kFwdFrags + kBwdFrags), dim3(64), 0, st, N->w, N->wf, N->wb);
}

static void ensure_cap(pyre_nrc* N, uint32_t n) {
  uint32_t need = (n + 15) / 16 * 16;
  if (need <= N->cap) return;
  dev_free(N->acts);
  dev_free(N->dz0);
  dev_free(N->dz1);
  dev_free(N->pred);
  dev_free(N->part);
  dev_free(N->loss_part);
  dev_free(N->feats);
  dev_free(N->targets);
  N->cap = need;
  N->acts = dev_alloc<__bf16>((size_t)need * (128 + 5 * 64));
  N->dz0 = dev_alloc<__bf16>((size_t)need * 128);
  N->dz1 = dev_alloc<__bf16>((size_t)need * 128);
  N->pred = dev_alloc<float>((size_t)need * 16);
  N->part = dev_alloc<float>((size_t)blocks(need, kDwChunk) * 128 * 64);
  N->loss_part = dev_alloc<float>(blocks(need, 256));
  N->feats = dev_alloc<float>((size_t)need * 13);
  N->targets = dev_alloc<float>((size_t)need * 3);
}

static size_t act_offset(int layer, uint32_t cap) {
  return layer == 0 ? 0 : (size_t)cap * (128 + (layer - 1) * 64);
}

// Forward + loss + backward into N->grad. feats/targets are device arrays.
// *out += scale * sum(part[0..n)) (one block).
__global__ void k_sum_into(const float* part, uint32_t n, float scale, float* out) {
  __shared__ float red[256];
  float s = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out += scale * red[0];
}

float nrc_gradients_dev(pyre_nrc* N, const float* feats, const float* targets, uint32_t n, hipStream_t st,
                        float* loss_dev) {
  require(n > 0, "empty training batch");
  ensure_cap(N, n);
  const uint32_t np = (n + 15) / 16 * 16;
  const uint32_t cap = N->cap;
  // pad rows: copy features (rows past n reuse query 0; their gradient is masked)
  hipLaunchKernelGGL(k_encode, dim3(blocks(n, 256)), dim3(256), 0, st, feats, n, N->bounds, N->acts);
  if (np > n)
    PYRE_HIP_CHECK(hipMemsetAsync(N->acts + (size_t)n * 128, 0, (size_t)(np - n) * 128 * sizeof(__bf16), st));
  const unsigned tb = blocks(np / 16, 4);
  for (int l = 0; l < kLayers - 1; ++l)
    hipLaunchKernelGGL((k_fwd_layer<true>), dim3(tb), dim3(256), 0, st, N->wf, l, N->acts + act_offset(l, cap), np,
                       N->acts + act_offset(l + 1, cap), (float*)nullptr);
  hipLaunchKernelGGL((k_fwd_layer<false>), dim3(tb), dim3(256), 0, st, N->wf, kLayers - 1,
                     N->acts + act_offset(kLayers - 1, cap), np, (__bf16*)nullptr, N->pred);
  hipLaunchKernelGGL(k_loss, dim3(blocks(np, 256)), dim3(256), 0, st, N->pred, targets, np, n, N->dz0, N->loss_part);
  __bf16* dz = N->dz0;
  __bf16* dzn = N->dz1;
  const uint32_t nch = blocks(np, kDwChunk);
  for (int l = kLayers - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_bwd_dw, dim3(nch), dim3(256), 0, st, l, N->acts + act_offset(l, cap), dz, np, N->part);
    const int size = kIn[l] * kOut[l];
    hipLaunchKernelGGL(k_dw_reduce, dim3(blocks(size, 256)), dim3(256), 0, st, N->part, nch, size,
                       N->grad + param_offset(l));
    if (l > 0) {
      hipLaunchKernelGGL(k_bwd_dx, dim3(tb), dim3(256), 0, st, N->wb, l, dz, N->acts + act_offset(l, cap), np, dzn);
      std::swap(dz, dzn);
    }
  }
  if (loss_dev) {
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, st, N->loss_part, blocks(np, 256), 1.0f / (3.0f * n),
                       loss_dev);
    return 0.0f;
  }
  std::vector<float> lp(blocks(np, 256));
  PYRE_HIP_CHECK(hipMemcpyAsync(lp.data(), N->loss_part, lp.size() * 4, hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  double s = 0;
  for (float x : lp) s += x;
  N->last_loss = (float)(s / (3.0 * n));
  return N->last_loss;
}

void nrc_adam_step(pyre_nrc* N, hipStream_t st) {
  N->step++;
  const float c1 = 1.0f - powf(N->prm.beta1, (float)N->step);
  const float c2 = 1.0f - powf(N->prm.beta2, (float)N->step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(kParams, 256)), dim3(256), 0, st, N->w, N->grad, N->m, N->v, kParams,
                     N->prm.learning_rate, N->prm.beta1, N->prm.beta2, N->prm.epsilon, c1, c2);
  nrc_refresh_frags(N, st);
}

void nrc_infer_dev(pyre_nrc* N, const float* feats, uint32_t n, float* rgb, hipStream_t st) {
  if (n == 0) return;
  unsigned grid = std::max(1u, std::min(blocks((n + 15) / 16, kInferWaves), (unsigned)N->ctx->cus * 2));
  hipLaunchKernelGGL(k_infer, dim3(grid), dim3(64 * kInferWaves), 0, st, N->wf, feats, (const uint32_t*)nullptr, n,
                     (const uint32_t*)nullptr, N->bounds, rgb);
}

void nrc_infer_dev_count(pyre_nrc* N, const float* feats, const uint32_t* rows, const uint32_t* n_dev, uint32_t cap,
                         float* rgb, hipStream_t st) {
  if (cap == 0) return;
  unsigned grid = std::max(1u, std::min(blocks((cap + 15) / 16, kInferWaves), (unsigned)N->ctx->cus * 2));
  hipLaunchKernelGGL(k_infer, dim3(grid), dim3(64 * kInferWaves), 0, st, N->wf, feats, rows, cap, n_dev, N->bounds,
                     rgb);
}

}  // namespace pyre

// ------------------------------------------------------------------- C ABI
namespace {
template <class F>
int nrc_guard(pyre_ctx* ctx, F&& f) {
  try {
    PYRE_HIP_CHECK(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return PYRE_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PYRE_EINVAL;
  }
}
}  // namespace

extern "C" {

int pyre_nrc_create(pyre_ctx* ctx, const pyre_nrc_params* p, pyre_nrc** out) {
  if (!ctx || !p || !out) return PYRE_EINVAL;
  *out = nullptr;
  pyre_nrc* N = new (std::nothrow) pyre_nrc();
  if (!N) return PYRE_ENOMEM;
  N->ctx = ctx;
  N->prm = *p;
  int rc = nrc_guard(ctx, [&] {
    for (int a = 0; a < 3; ++a) {
      require(p->bounds_max[a] > p->bounds_min[a], "nrc bounds must have positive extent");
      N->bounds.lo[a] = p->bounds_min[a];
      N->bounds.inv_ext[a] = 1.0f / (p->bounds_max[a] - p->bounds_min[a]);
    }
    // He-uniform initialisation from xoshiro256++ (zero padding rows of layer
    // 0 past the constant slot, zero non-RGB output columns)
    std::vector<float> w(kParams);
    Rng r = rng_seed(p->seed);
    for (int l = 0; l < kLayers; ++l) {
      float lim = sqrtf(6.0f / (float)kIn[l]);
      for (int i = 0; i < kIn[l]; ++i)
        for (int o = 0; o < kOut[l]; ++o) {
          float u = rng_f32(r) * 2.0f - 1.0f;
          bool zero = (l == kLayers - 1 && o >= 3) || (l == 0 && i > 108);
          w[param_offset(l) + i * kOut[l] + o] = zero ? 0.0f : u * lim;
        }
    }
    hipStream_t st = ctx->stream;
    N->w = dev_alloc<float>(kParams);
    N->m = dev_alloc<float>(kParams);
    N->v = dev_alloc<float>(kParams);
    N->grad = dev_alloc<float>(kParams);
    N->wf = dev_alloc<bf16x8>((size_t)kFwdFrags * 64);
    N->wb = dev_alloc<bf16x8>((size_t)kBwdFrags * 64);
    PYRE_HIP_CHECK(hipMemcpyAsync(N->w, w.data(), kParams * 4, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemsetAsync(N->m, 0, kParams * 4, st));
    PYRE_HIP_CHECK(hipMemsetAsync(N->v, 0, kParams * 4, st));
    nrc_refresh_frags(N, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
  });
  if (rc != PYRE_OK) {
    pyre_nrc_destroy(N);
    return rc;
  }
  *out = N;
  return PYRE_OK;
}

void pyre_nrc_destroy(pyre_nrc* N) {
  if (!N) return;
  (void)hipSetDevice(N->ctx->device);
  dev_free(N->w);
  dev_free(N->m);
  dev_free(N->v);
  dev_free(N->grad);
  dev_free(N->wf);
  dev_free(N->wb);
  dev_free(N->acts);
  dev_free(N->dz0);
  dev_free(N->dz1);
  dev_free(N->pred);
  dev_free(N->part);
  dev_free(N->loss_part);
  dev_free(N->feats);
  dev_free(N->targets);
  delete N;
}

size_t pyre_nrc_param_count(void) { return (size_t)kParams; }

int pyre_nrc_get_weights(pyre_ctx* ctx, pyre_nrc* N, float* out) {
  if (!ctx || !N || !out) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    PYRE_HIP_CHECK(hipMemcpyAsync(out, N->w, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pyre_nrc_set_weights(pyre_ctx* ctx, pyre_nrc* N, const float* w) {
  if (!ctx || !N || !w) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    PYRE_HIP_CHECK(hipMemcpyAsync(N->w, w, kParams * 4, hipMemcpyHostToDevice, ctx->stream));
    PYRE_HIP_CHECK(hipMemsetAsync(N->m, 0, kParams * 4, ctx->stream));
    PYRE_HIP_CHECK(hipMemsetAsync(N->v, 0, kParams * 4, ctx->stream));
    N->step = 0;
    nrc_refresh_frags(N, ctx->stream);
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pyre_nrc_infer(pyre_ctx* ctx, pyre_nrc* N, const float* feats, size_t n, float* rgb) {
  if (!ctx || !N || (n && (!feats || !rgb))) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    if (!n) return;
    hipStream_t st = ctx->stream;
    float* f = dev_alloc<float>(13 * n);
    float* o = dev_alloc<float>(3 * n);
    PYRE_HIP_CHECK(hipMemcpyAsync(f, feats, 13 * 4 * n, hipMemcpyHostToDevice, st));
    nrc_infer_dev(N, f, (uint32_t)n, o, st);
    PYRE_HIP_CHECK(hipMemcpyAsync(rgb, o, 3 * 4 * n, hipMemcpyDeviceToHost, st));
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    dev_free(f);
    dev_free(o);
  });
}

int pyre_nrc_gradients(pyre_ctx* ctx, pyre_nrc* N, const float* feats, const float* targets, size_t n, float* grad,
                       float* loss) {
  if (!ctx || !N || !feats || !targets || !n) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(n < (1u << 26), "training batch too large");
    ensure_cap(N, (uint32_t)n);
    hipStream_t st = ctx->stream;
    PYRE_HIP_CHECK(hipMemcpyAsync(N->feats, feats, 13 * 4 * n, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemcpyAsync(N->targets, targets, 3 * 4 * n, hipMemcpyHostToDevice, st));
    float l = nrc_gradients_dev(N, N->feats, N->targets, (uint32_t)n, st, nullptr);
    if (loss) *loss = l;
    if (grad) {
      PYRE_HIP_CHECK(hipMemcpyAsync(grad, N->grad, kParams * 4, hipMemcpyDeviceToHost, st));
      PYRE_HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

int pyre_nrc_stats(pyre_ctx* ctx, pyre_nrc* N, uint64_t* steps, float* last_loss) {
  if (!ctx || !N) return PYRE_EINVAL;
  if (steps) *steps = N->step;
  if (last_loss) *last_loss = N->last_loss;
  return PYRE_OK;
}

int pyre_nrc_train_step(pyre_ctx* ctx, pyre_nrc* N, const float* feats, const float* targets, size_t n, float* loss) {
  if (!ctx || !N || !feats || !targets || !n) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(n < (1u << 26), "training batch too large");
    ensure_cap(N, (uint32_t)n);
    hipStream_t st = ctx->stream;
    PYRE_HIP_CHECK(hipMemcpyAsync(N->feats, feats, 13 * 4 * n, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemcpyAsync(N->targets, targets, 3 * 4 * n, hipMemcpyHostToDevice, st));
    float l = nrc_gradients_dev(N, N->feats, N->targets, (uint32_t)n, st, nullptr);
    nrc_adam_step(N, st);
    PYRE_HIP_CHECK(hipStreamSynchronize(st));
    if (loss) *loss = l;
  });
}

}  // extern "C"
