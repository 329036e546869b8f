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
// Features 2i and 2i + 1 below 72 are sin and cos of one argument: one
// sincospi per pair.
__device__ __forceinline__ bf16x8 encode_frag(const Query& q, int kb, int g) {
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    const int f = 32 * kb + 8 * g + j;
    if (f < 72) {
      const int a = f / 24, k = (f % 24) / 2;
      const float pa = a == 0 ? q.px : (a == 1 ? q.py : q.pz);
      float sv, cv;
      sincospif(ldexpf(pa, k), &sv, &cv);
      v[j] = to_bf16(sv);
      v[j + 1] = to_bf16(cv);
    } else {
      v[j] = to_bf16(encode_feature(f, q));
      v[j + 1] = to_bf16(encode_feature(f + 1, q));
    }
  }
  return v;
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

// dst[0..N) = src[0..N) by a block of THREADS threads: every load is issued
// before the first LDS store, so the copy costs one memory latency, not one
// per element a thread moves.
template <int N, int THREADS>
__device__ __forceinline__ void stage_frags(bf16x8* dst, const bf16x8* __restrict__ src) {
  constexpr int kIt = (N + THREADS - 1) / THREADS;
  bf16x8 v[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = (int)threadIdx.x + it * THREADS;
    if (N % THREADS == 0 || i < N) v[it] = src[i];
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = (int)threadIdx.x + it * THREADS;
    if (N % THREADS == 0 || i < N) dst[i] = v[it];
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
  stage_frags<kFwdFrags * 64, 64 * kInferWaves>(sw, wfrag);
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
// Weight gradients are summed over chunks of kDwChunk samples per block.
constexpr int kDwChunk = 128;

// Four launches per step: fused encode + forward + loss gradient, fused
// backward dX through all layers, weight gradients of all layers, and their
// deterministic reduction fused with the Adam update.
constexpr int kTrainWaves = 4;

// Size of a training batch: n given by the host, or, with count, the samples
// [off, off + n) of the *count (at most cap) a device stage produced; none
// when fewer than 16 remain, so a step can be queued before its size is known.
struct BatchN {
  const uint32_t* count;
  uint32_t cap, off, n;
};
__device__ __forceinline__ uint32_t batch_size(const BatchN& b) {
  if (!b.count) return b.n;
  const uint32_t avail = min(*b.count, b.cap);
  return avail >= b.off + 16u ? min(b.n, avail - b.off) : 0u;
}

// Encode + forward of every layer + relative-L2 loss gradient for tiles of 16
// samples. Writes each layer's input (bf16, sample-major), the loss gradient
// dP (16 wide, zero past RGB and for padding rows q >= n_real) and per-block
// loss sums; counts the Adam step in *step when given and the batch is not
// empty.
__global__ __launch_bounds__(64 * kTrainWaves) void k_train_fwd(const bf16x8* __restrict__ wfrag,
                                                                const float* __restrict__ feats,
                                                                const float* __restrict__ targets, BatchN bn,
                                                                NrcBounds b, __bf16* acts, uint32_t cap, __bf16* dp,
                                                                float* loss_part, uint32_t* step) {
  const uint32_t n_real = batch_size(bn), np = (n_real + 15u) / 16u * 16u;
  if (step && n_real && blockIdx.x == 0 && threadIdx.x == 0) *step += 1u;
  __shared__ bf16x8 sw[kFwdFrags * 64];
  __shared__ __attribute__((aligned(16))) __bf16 sact[kTrainWaves][16][64 + 8];
  __shared__ float red[kTrainWaves];
  stage_frags<kFwdFrags * 64, 64 * kTrainWaves>(sw, wfrag);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, r = lane & 15;
  const uint32_t ntiles = np / 16;
  float loss = 0.0f;
  for (uint32_t tile = blockIdx.x * kTrainWaves + wave; tile < ntiles; tile += gridDim.x * kTrainWaves) {
    const uint32_t q = tile * 16 + r;
    const uint32_t row = q < n_real ? q : n_real - 1;  // padding rows repeat a sample; their gradient is zero
    Query qq = load_query(feats + 13 * (size_t)row, b);
    f32x4 acc[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) acc[ot] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      bf16x8 bx = encode_frag(qq, kb, g);
      *(bf16x8*)(acts + (size_t)q * 128 + 32 * kb + 8 * g) = bx;
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) acc[ot] = mfma(sw[(fwd_frag_offset(0) + ot * 4 + kb) * 64 + lane], bx, acc[ot]);
    }
#pragma unroll
    for (int l = 1; l < kLayers; ++l) {
      __bf16* xl = acts + (size_t)cap * (128 + (l - 1) * 64) + (size_t)q * 64;
#pragma unroll
      for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int i = 0; i < 4; ++i) sact[wave][r][16 * ot + 4 * g + i] = to_bf16(fmaxf(acc[ot][i], 0.0f));
      wave_lds_sync();
      // B-operand layout: lane (g, r) holds features 32 kb + 8 g .. + 7 of
      // sample q, which is also one 16-byte piece of its activation row
      bf16x8 bx[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        bx[kb] = *(const bf16x8*)&sact[wave][r][32 * kb + 8 * g];
        *(bf16x8*)(xl + 32 * kb + 8 * g) = bx[kb];
      }
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
    // outputs 4g + i of sample q; relative L2 on RGB: L = (p - t)^2 / (p^2 + 0.01)
    // with p^2 held constant in the gradient (the stop-gradient form)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = 4 * g + i;
      float gval = 0.0f;
      if (c < 3 && q < n_real) {
        const float p = acc[0][i], t = targets[3 * (size_t)q + c];
        const float denom = p * p + 0.01f;
        const float d = p - t;
        loss += d * d / denom;
        gval = 2.0f * d / denom / (3.0f * (float)n_real);
      }
      dp[(size_t)q * 16 + c] = to_bf16(gval);
    }
  }
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  if (lane == 0) red[wave] = loss;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < kTrainWaves; ++w) t += red[w];
    loss_part[blockIdx.x] = t;
  }
}

// Backward dX through layers 5..1 for tiles of 16 samples:
// dZ_{l-1} = (dZ_l W_l^T) * (X_l > 0), every dZ_l kept for the weight
// gradients (dz + dz_offset(l)).
__host__ __device__ constexpr size_t dz_offset(int l, uint32_t cap) { return (size_t)cap * 64 * (size_t)l; }

__global__ __launch_bounds__(64 * kTrainWaves) void k_train_bwd(const bf16x8* __restrict__ bfrag,
                                                                const __bf16* __restrict__ acts, BatchN bn,
                                                                uint32_t cap, __bf16* dz) {
  const uint32_t np = (batch_size(bn) + 15u) / 16u * 16u;
  constexpr int kUsed = kBwdFrags - bwd_frags(0);  // layers 1..5
  __shared__ bf16x8 sw[kUsed * 64];
  __shared__ __attribute__((aligned(16))) __bf16 sdz[kTrainWaves][16][64 + 8];
  stage_frags<kUsed * 64, 64 * kTrainWaves>(sw, bfrag + bwd_frag_offset(1) * 64);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, r = lane & 15;
  const uint32_t ntiles = np / 16;
  for (uint32_t tile = blockIdx.x * kTrainWaves + wave; tile < ntiles; tile += gridDim.x * kTrainWaves) {
    const uint32_t q = tile * 16 + r;
    // dZ_5 = dP (16 wide, zero-padded to the K block of 32)
    bf16x8 bz[2];
    {
      const __bf16* d5 = dz + dz_offset(5, cap) + (size_t)q * 16;
      if (g < 2) {
        bz[0] = *(const bf16x8*)(d5 + 8 * g);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) bz[0][j] = to_bf16(0.0f);
      }
    }
#pragma unroll
    for (int l = kLayers - 1; l >= 1; --l) {
      const int nkb = (kOut[l] + 31) / 32;
      const __bf16* xl = acts + (size_t)cap * (128 + (l - 1) * 64) + (size_t)q * 64;
      __bf16* dzo = dz + dz_offset(l - 1, cap) + (size_t)q * 64;
#pragma unroll
      for (int it = 0; it < 4; ++it) {  // in = 64 for layers 1..5
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          if (kb < nkb) acc = mfma(sw[(bwd_frag_offset(l) - bwd_frag_offset(1) + it * nkb + kb) * 64 + lane], bz[kb], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) sdz[wave][r][16 * it + 4 * g + i] = to_bf16(acc[i]);
      }
      wave_lds_sync();
      // ReLU mask in the B-operand layout (features 32 kb + 8 g .. + 7),
      // 16-byte loads of X_l and stores of dZ_{l-1}
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        const bf16x8 raw = *(const bf16x8*)&sdz[wave][r][32 * kb + 8 * g];
        const bf16x8 xv = *(const bf16x8*)(xl + 32 * kb + 8 * g);
#pragma unroll
        for (int j = 0; j < 8; ++j) bz[kb][j] = from_bf16(xv[j]) > 0.0f ? raw[j] : to_bf16(0.0f);
        *(bf16x8*)(dzo + 32 * kb + 8 * g) = bz[kb];
      }
      wave_lds_sync();
    }
  }
}

// Weight gradients of every layer: block (p, layer) sums X_l^T dZ_l over the
// chunks p, p + P, ... of kDwChunk samples (P = gridDim.x) in registers and
// writes one partial. Each chunk is staged feature-major in LDS, so both MFMA
// operands (A = X^T tile, B = dZ tile; K = samples) are single 16-byte reads;
// wave w owns the 16x16 output tiles w, w + 4, ... of the layer.
constexpr int kDwTiles = 8;               // tiles per wave: 128 x 64 / 256 / 4
constexpr int kDwStride = kDwChunk + 8;   // bf16 per LDS row: 16-byte aligned, conflict-free 16-byte reads
static_assert(param_offset(1) % 64 == 0 && param_offset(2) % 64 == 0 && param_offset(5) % 64 == 0 &&
                  kParams % 64 == 0,
              "parameter blocks of 64 stay within one layer");

// One chunk's samples held in registers between its global loads and its
// feature-major LDS copy (dst[f * kDwStride + s] = src[(q0 + s) * width + f],
// zero past np). Work items are (sample pair, 8-feature group), sample pair
// fastest, so each LDS store is one 32-bit word holding two samples.
constexpr int kPairs = kDwChunk / 2;
template <int W>
struct ChunkRegs {
  static constexpr int kTotal = kPairs * (W / 8);
  static constexpr int kItems = (kTotal + 255) / 256;  // per thread (blocks of 256)
  uint4 a[kItems], c[kItems];
  __device__ __forceinline__ void load(const __bf16* __restrict__ src, uint32_t q0, uint32_t np) {
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int i = threadIdx.x + 256 * it;
      const int sp = i % kPairs, f8 = i / kPairs;
      const uint32_t s = 2 * sp;
      if (kTotal % 256 != 0 && i >= kTotal) break;
      a[it] = q0 + s < np ? *(const uint4*)(src + (size_t)(q0 + s) * W + 8 * f8) : make_uint4(0, 0, 0, 0);
      c[it] = q0 + s + 1 < np ? *(const uint4*)(src + (size_t)(q0 + s + 1) * W + 8 * f8) : make_uint4(0, 0, 0, 0);
    }
  }
  __device__ __forceinline__ void store(__bf16* dst) const {
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int i = threadIdx.x + 256 * it;
      const int sp = i % kPairs, f8 = i / kPairs;
      const uint32_t s = 2 * sp;
      if (kTotal % 256 != 0 && i >= kTotal) break;
      const uint32_t wa[4] = {a[it].x, a[it].y, a[it].z, a[it].w}, wc[4] = {c[it].x, c[it].y, c[it].z, c[it].w};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t lo = (wa[j >> 1] >> (16 * (j & 1))) & 0xffffu;
        const uint32_t hi = (wc[j >> 1] >> (16 * (j & 1))) & 0xffffu;
        d[((8 * f8 + j) * kDwStride + s) >> 1] = lo | (hi << 16);
      }
    }
  }
};

// Block (p, layer) of k_train_dw for a layer of IN inputs and OUT outputs:
// chunk c + P's loads are in flight while chunk c's MFMAs run.
template <int IN, int OUT>
__device__ __forceinline__ void dw_layer(const __bf16* __restrict__ xl, const __bf16* __restrict__ dzl, uint32_t np,
                                         __bf16* sx, __bf16* sz, f32x4 (&acc)[kDwTiles]) {
  const uint32_t nch = (np + kDwChunk - 1) / kDwChunk;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  constexpr int kNot = OUT / 16, kTiles = (IN / 16) * kNot;
  ChunkRegs<IN> rx;
  ChunkRegs<OUT> rz;
  uint32_t c = blockIdx.x;
  if (c < nch) {
    rx.load(xl, c * kDwChunk, np);
    rz.load(dzl, c * kDwChunk, np);
  }
  for (; c < nch; c += gridDim.x) {
    __syncthreads();  // the previous chunk's LDS reads are done
    rx.store(sx);
    rz.store(sz);
    __syncthreads();
    if (c + gridDim.x < nch) {
      rx.load(xl, (c + gridDim.x) * kDwChunk, np);
      rz.load(dzl, (c + gridDim.x) * kDwChunk, np);
    }
#pragma unroll
    for (int k = 0; k < kDwTiles; ++k) {
      const int t = wave + 4 * k;
      if (t < kTiles) {
        const int it = t / kNot, ot = t % kNot;
        const __bf16* ra = sx + (16 * it + r) * kDwStride + 8 * g;
        const __bf16* rb = sz + (16 * ot + r) * kDwStride + 8 * g;
#pragma unroll
        for (int kb = 0; kb < kDwChunk / 32; ++kb)
          acc[k] = mfma(*(const bf16x8*)(ra + 32 * kb), *(const bf16x8*)(rb + 32 * kb), acc[k]);
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_train_dw(const __bf16* __restrict__ acts, const __bf16* __restrict__ dz,
                                                  BatchN bn, uint32_t cap, float* partial) {
  const uint32_t np = (batch_size(bn) + 15u) / 16u * 16u;
  __shared__ __attribute__((aligned(16))) __bf16 sx[128 * kDwStride];
  __shared__ __attribute__((aligned(16))) __bf16 sz[64 * kDwStride];
  const int layer = blockIdx.y;
  const int in = kIn[layer], out = kOut[layer];
  const __bf16* xl = acts + (layer == 0 ? 0 : (size_t)cap * (128 + (layer - 1) * 64));
  const __bf16* dzl = dz + dz_offset(layer, cap);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const int not_ = out / 16, ntile = (in / 16) * not_;
  f32x4 acc[kDwTiles];
#pragma unroll
  for (int k = 0; k < kDwTiles; ++k) acc[k] = f32x4{0, 0, 0, 0};
  if (layer == 0)
    dw_layer<128, 64>(xl, dzl, np, sx, sz, acc);
  else if (layer == kLayers - 1)
    dw_layer<64, 16>(xl, dzl, np, sx, sz, acc);
  else
    dw_layer<64, 64>(xl, dzl, np, sx, sz, acc);
  float* dst = partial + (size_t)gridDim.x * param_offset(layer) + (size_t)blockIdx.x * in * out;
#pragma unroll
  for (int k = 0; k < kDwTiles; ++k) {
    const int t = wave + 4 * k;
    if (t < ntile) {
      const int it = t / not_, ot = t % not_;
#pragma unroll
      for (int i = 0; i < 4; ++i) dst[(size_t)(16 * it + 4 * g + i) * out + 16 * ot + r] = acc[k][i];
    }
  }
}

// grad[i] = sum of the P partials: block of 64 parameters (never straddling a
// layer: every layer size is a multiple of 64), wave w sums partials w, w + 4,
// ... and the four wave sums are added in wave order (deterministic).
__global__ __launch_bounds__(256) void k_train_reduce(const float* partial, uint32_t P, float* grad) {
  __shared__ float red[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 64 + lane;
  int l = 0;
  while (l + 1 < kLayers && blockIdx.x * 64 >= param_offset(l + 1)) ++l;
  const int size = kIn[l] * kOut[l], j = i - param_offset(l);
  const float* p = partial + (size_t)P * param_offset(l) + j;
  float s = 0.0f;
  for (uint32_t c = wave; c < P; c += 4) s += p[(size_t)c * size];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0) grad[i] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Adam hyper-parameters with a step's bias corrections c1 = 1 - b1^t,
// c2 = 1 - b2^t.
struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;
};

__device__ __forceinline__ float adam_update(float& m, float& v, float w, float gr, const AdamArgs& a) {
  m = a.b1 * m + (1.0f - a.b1) * gr;
  v = a.b2 * v + (1.0f - a.b2) * gr * gr;
  const float mh = m / a.c1, vh = v / a.c2;
  return w - a.lr * mh / (sqrtf(vh) + a.eps);
}

// The tail of a training step in one launch: grad = sum of the P partials
// (in the order of k_train_reduce), Adam on the FP32 master weights, and the
// new weights rounded to bf16 straight into their forward and backward
// fragment slots (the inverse of k_make_frags' gather; the backward
// fragments' zero padding is never written). The non-RGB output columns and
// the padding rows of layer 0 have zero gradient and stay fixed. The step
// number t is *step (counted by k_train_fwd); an empty batch changes nothing.
// Block 0 also adds the batch's mean loss to *loss_dev and counts the step
// in *steps_dev when they are given.
__global__ __launch_bounds__(256) void k_train_update(const float* partial, uint32_t P, float* grad, float* w,
                                                      float* m, float* v, AdamArgs a, const uint32_t* step,
                                                      BatchN bn, bf16x8* wf, bf16x8* wb, const float* loss_part,
                                                      uint32_t nlp, float* loss_dev, uint32_t* steps_dev) {
  __shared__ float red[4][64];
  const uint32_t n = batch_size(bn);
  if (!n) return;
  const float t = (float)*step;
  a.c1 = 1.0f - powf(a.b1, t);
  a.c2 = 1.0f - powf(a.b2, t);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 64 + lane;
  int l = 0;
  while (l + 1 < kLayers && blockIdx.x * 64 >= param_offset(l + 1)) ++l;
  const int in = kIn[l], out = kOut[l], j = i - param_offset(l);
  const float* p = partial + (size_t)P * param_offset(l) + j;
  float s = 0.0f;
  for (uint32_t c = wave; c < P; c += 4) s += p[(size_t)c * (in * out)];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0) {
    const float gr = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    grad[i] = gr;
    float mm = m[i], vv = v[i];
    const float wn = adam_update(mm, vv, w[i], gr, a);
    m[i] = mm;
    v[i] = vv;
    w[i] = wn;
    const __bf16 hb = to_bf16(wn);
    const int ji = j / out, o = j % out;  // W_l[ji][o]
    {
      const int frag = fwd_frag_offset(l) + (o / 16) * (in / 32) + ji / 32;
      const int ln = ((ji % 32) / 8) * 16 + o % 16;
      reinterpret_cast<__bf16*>(wf)[((size_t)frag * 64 + ln) * 8 + ji % 8] = hb;
    }
    {
      const int frag = bwd_frag_offset(l) + (ji / 16) * ((out + 31) / 32) + o / 32;
      const int ln = ((o % 32) / 8) * 16 + ji % 16;
      reinterpret_cast<__bf16*>(wb)[((size_t)frag * 64 + ln) * 8 + o % 8] = hb;
    }
  }
  if (loss_dev && blockIdx.x == 0) {
    __shared__ float lred[256];
    float t = 0.0f;
    for (uint32_t k = threadIdx.x; k < nlp; k += 256) t += loss_part[k];
    lred[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) lred[threadIdx.x] += lred[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) *loss_dev += lred[0] / (3.0f * (float)n);
  }
  if (steps_dev && blockIdx.x == 0 && threadIdx.x == 0) *steps_dev += 1u;
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Weight-gradient partials per layer: at least four chunks per block, at most
// 256 partials (26 MB): large batches get more blocks for the (memory-bound)
// weight gradients at the price of a longer reduction.
inline unsigned dw_partials(uint32_t np) { return std::max(1u, std::min(256u, blocks(np, 4 * kDwChunk))); }

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
  uint32_t* d_step = nullptr;  // Adam steps taken (device counter)
  float last_loss = 0.0f;
  // training scratch, sized for the largest batch seen
  uint32_t cap = 0;
  __bf16* acts = nullptr;  // layer inputs: 128 + 5 * 64 features per query
  __bf16* dz = nullptr;  // every layer's output gradient (dz_offset)
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
  dev_free(N->dz);
  dev_free(N->part);
  dev_free(N->loss_part);
  dev_free(N->feats);
  dev_free(N->targets);
  N->cap = need;
  N->acts = dev_alloc<__bf16>((size_t)need * (128 + 5 * 64));
  N->dz = dev_alloc<__bf16>((size_t)need * (5 * 64 + 16));
  N->part = dev_alloc<float>((size_t)dw_partials(need) * kParams);
  N->loss_part = dev_alloc<float>(std::max<size_t>(blocks(need, 64), (size_t)std::max(1, N->ctx->cus) * 2));
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

// Forward + loss + backward + weight-gradient partials of one batch of at
// most bn.n samples (feats/targets are device arrays); step counts the Adam
// step on the device. Returns the number of per-block loss sums in
// N->loss_part; the partials' count is dw_partials of the padded bn.n.
static unsigned train_pass(pyre_nrc* N, const float* feats, const float* targets, const BatchN& bn, uint32_t* step,
                           hipStream_t st) {
  require(bn.n > 0, "empty training batch");
  ensure_cap(N, bn.n);
  const uint32_t np = (bn.n + 15) / 16 * 16;
  const uint32_t cap = N->cap;
  const unsigned tiles = np / 16;
  const unsigned grid = std::max(1u, std::min(blocks(tiles, kTrainWaves), (unsigned)std::max(1, N->ctx->cus) * 2));
  hipLaunchKernelGGL(k_train_fwd, dim3(grid), dim3(64 * kTrainWaves), 0, st, N->wf, feats, targets, bn, N->bounds,
                     N->acts, cap, N->dz + dz_offset(5, cap), N->loss_part, step);
  hipLaunchKernelGGL(k_train_bwd, dim3(grid), dim3(64 * kTrainWaves), 0, st, N->wb, N->acts, bn, cap, N->dz);
  hipLaunchKernelGGL(k_train_dw, dim3(dw_partials(np), kLayers), dim3(256), 0, st, N->acts, N->dz, bn, cap, N->part);
  return grid;
}

// Mean loss of the batch from the per-block sums (synchronises the stream).
static float read_loss(pyre_nrc* N, unsigned nlp, uint32_t n, hipStream_t st) {
  std::vector<float> lp(nlp);
  PYRE_HIP_CHECK(hipMemcpyAsync(lp.data(), N->loss_part, lp.size() * 4, hipMemcpyDeviceToHost, st));
  PYRE_HIP_CHECK(hipStreamSynchronize(st));
  double sum = 0;
  for (float x : lp) sum += x;
  N->last_loss = (float)(sum / (3.0 * n));
  return N->last_loss;
}

static AdamArgs adam_args(const pyre_nrc* N) {
  return AdamArgs{N->prm.learning_rate, N->prm.beta1, N->prm.beta2, N->prm.epsilon, 0.0f, 0.0f};
}

float nrc_gradients_dev(pyre_nrc* N, const float* feats, const float* targets, uint32_t n, hipStream_t st,
                        float* loss_dev) {
  const BatchN bn{nullptr, 0, 0, n};
  const unsigned nlp = train_pass(N, feats, targets, bn, nullptr, st);
  const uint32_t np = (n + 15) / 16 * 16;
  hipLaunchKernelGGL(k_train_reduce, dim3(kParams / 64), dim3(256), 0, st, N->part, dw_partials(np), N->grad);
  if (loss_dev) {
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, st, N->loss_part, nlp, 1.0f / (3.0f * n), loss_dev);
    return 0.0f;
  }
  return read_loss(N, nlp, n, st);
}

float nrc_train_step_dev(pyre_nrc* N, const float* feats, const float* targets, uint32_t n, hipStream_t st,
                         float* loss_dev, bool read_back) {
  const BatchN bn{nullptr, 0, 0, n};
  const unsigned nlp = train_pass(N, feats, targets, bn, N->d_step, st);
  const uint32_t np = (n + 15) / 16 * 16;
  hipLaunchKernelGGL(k_train_update, dim3(kParams / 64), dim3(256), 0, st, N->part, dw_partials(np), N->grad, N->w,
                     N->m, N->v, adam_args(N), (const uint32_t*)N->d_step, bn, N->wf, N->wb, N->loss_part, nlp,
                     loss_dev, (uint32_t*)nullptr);
  return loss_dev || !read_back ? 0.0f : read_loss(N, nlp, n, st);
}

void nrc_train_step_counted(pyre_nrc* N, const float* feats, const float* targets, const uint32_t* count,
                            uint32_t cap, uint32_t off, uint32_t batch, hipStream_t st, float* loss_dev,
                            uint32_t* steps_dev) {
  const BatchN bn{count, cap, off, batch};
  const unsigned nlp = train_pass(N, feats + 13 * (size_t)off, targets + 3 * (size_t)off, bn, N->d_step, st);
  const uint32_t np = (batch + 15) / 16 * 16;
  hipLaunchKernelGGL(k_train_update, dim3(kParams / 64), dim3(256), 0, st, N->part, dw_partials(np), N->grad, N->w,
                     N->m, N->v, adam_args(N), (const uint32_t*)N->d_step, bn, N->wf, N->wb, N->loss_part, nlp,
                     loss_dev, steps_dev);
}

void nrc_set_last_loss(pyre_nrc* N, float loss) { N->last_loss = loss; }

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
    N->d_step = dev_alloc<uint32_t>(1);
    PYRE_HIP_CHECK(hipMemcpyAsync(N->w, w.data(), kParams * 4, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemsetAsync(N->m, 0, kParams * 4, st));
    PYRE_HIP_CHECK(hipMemsetAsync(N->v, 0, kParams * 4, st));
    PYRE_HIP_CHECK(hipMemsetAsync(N->d_step, 0, 4, st));
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
  dev_free(N->d_step);
  dev_free(N->acts);
  dev_free(N->dz);
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
    PYRE_HIP_CHECK(hipMemsetAsync(N->d_step, 0, 4, ctx->stream));
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

int pyre_nrc_reserve(pyre_ctx* ctx, pyre_nrc* N, size_t max_batch) {
  if (!ctx || !N || !max_batch) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(max_batch < (1u << 26), "training batch too large");
    ensure_cap(N, (uint32_t)max_batch);
  });
}

int pyre_nrc_stats(pyre_ctx* ctx, pyre_nrc* N, uint64_t* steps, float* last_loss) {
  if (!ctx || !N) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    uint32_t t = 0;
    PYRE_HIP_CHECK(hipMemcpyAsync(&t, N->d_step, 4, hipMemcpyDeviceToHost, ctx->stream));
    PYRE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (steps) *steps = t;
    if (last_loss) *last_loss = N->last_loss;
  });
}

int pyre_nrc_infer_dev(pyre_ctx* ctx, pyre_nrc* N, const float* feats, size_t n, float* rgb, void* stream) {
  if (!ctx || !N || (n && (!feats || !rgb))) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(n < (1ull << 31), "too many queries");
    if (n) nrc_infer_dev(N, feats, (uint32_t)n, rgb, stream ? (hipStream_t)stream : ctx->stream);
  });
}

int pyre_nrc_train_step_dev(pyre_ctx* ctx, pyre_nrc* N, const float* feats, const float* targets, size_t n,
                            float* loss, void* stream) {
  if (!ctx || !N || !feats || !targets || !n) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(n < (1u << 26), "training batch too large");
    nrc_train_step_dev(N, feats, targets, (uint32_t)n, stream ? (hipStream_t)stream : ctx->stream, loss, false);
  });
}

int pyre_nrc_train_step(pyre_ctx* ctx, pyre_nrc* N, const float* feats, const float* targets, size_t n, float* loss) {
  if (!ctx || !N || !feats || !targets || !n) return PYRE_EINVAL;
  return nrc_guard(ctx, [&] {
    require(n < (1u << 26), "training batch too large");
    ensure_cap(N, (uint32_t)n);
    hipStream_t st = ctx->stream;
    PYRE_HIP_CHECK(hipMemcpyAsync(N->feats, feats, 13 * 4 * n, hipMemcpyHostToDevice, st));
    PYRE_HIP_CHECK(hipMemcpyAsync(N->targets, targets, 3 * 4 * n, hipMemcpyHostToDevice, st));
    float l = nrc_train_step_dev(N, N->feats, N->targets, (uint32_t)n, st, nullptr);
    if (loss) *loss = l;
  });
}

}  // extern "C"
