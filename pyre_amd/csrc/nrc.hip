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
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(8 * sizeof(float)))) float f32x8;

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

// A query's 13 input floats, fetched ahead of their use.
struct RawQuery {
  float f[13];
};

__device__ __forceinline__ RawQuery fetch_query(const float* f) {
  RawQuery r;
#pragma unroll
  for (int k = 0; k < 13; ++k) r.f[k] = f[k];
  return r;
}

__device__ __forceinline__ Query query_of(const RawQuery& r, const NrcBounds& b) {
  Query q;
  q.px = (r.f[0] - b.lo[0]) * b.inv_ext[0];
  q.py = (r.f[1] - b.lo[1]) * b.inv_ext[1];
  q.pz = (r.f[2] - b.lo[2]) * b.inv_ext[2];
  q.dx = r.f[3];
  q.dy = r.f[4];
  q.dz = r.f[5];
  q.nx = r.f[6];
  q.ny = r.f[7];
  q.nz = r.f[8];
  q.rough = r.f[9];
  q.ar = r.f[10];
  q.ag = r.f[11];
  q.ab = r.f[12];
  return q;
}

__device__ __forceinline__ Query load_query(const float* f, const NrcBounds& b) { return query_of(fetch_query(f), b); }

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
// Features 2i and 2i + 1 below 72 are sin and cos of one argument, and the 8
// features of a lane below 72 are the frequencies k0 .. k0 + 3 of one axis
// (k0 = 0, 4 or 8): one sincospi at 2^k0 p, then three angle doublings
// (sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x), whose rounding error,
// below 1e-6, is far under the bf16 step the features are stored with.
__device__ __forceinline__ bf16x8 encode_frag(const Query& q, int kb, int g) {
  f32x8 v;
  const int f0 = 32 * kb + 8 * g;
  if (f0 < 72) {
    const int a = f0 / 24, k0 = (f0 % 24) / 2;
    const float pa = a == 0 ? q.px : (a == 1 ? q.py : q.pz);
    float s, c;
    sincospif(ldexpf(pa, k0), &s, &c);
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      v[j] = s;
      v[j + 1] = c;
      const float s2 = s + s;
      const float sn = s2 * c;
      c = fmaf(-s2, s, 1.0f);
      s = sn;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = encode_feature(f0 + j, q);
  }
  return __builtin_convertvector(v, bf16x8);  // round to nearest even, two values per instruction
}

typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Two floats rounded to bf16 (nearest even) in one v_cvt_pk_bf16_f32, lo in
// bits 0..15 (converting pairs: from wider vectors the compiler rounds one
// value per instruction and merges the halves).
__device__ __forceinline__ unsigned cvt_pk_bf16(float lo, float hi) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

// ReLU of 4 accumulators rounded to bf16: the maximum with 0 of the rounded
// values read as signed 16-bit integers (negative values and -0 become +0),
// one packed integer instruction per pair.
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 a) {
  unsigned out[2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
    out[h] = __builtin_bit_cast(
        unsigned, __builtin_elementwise_max(__builtin_bit_cast(s16x2, cvt_pk_bf16(a[2 * h], a[2 * h + 1])), (s16x2)0));
  return __builtin_bit_cast(bf16x4, out);
}

// Gradients through a ReLU: bf16(a) where the layer's stored output x is
// positive, else 0. x comes from relu_bf16, so its bit patterns are +0 or a
// positive number: min(bits, 1) per 16-bit half is the 0/1 mask, applied by a
// 16-bit multiply -- two packed instructions per pair of values. (The 1 goes
// through an empty asm statement, which emits nothing: a multiply by a value
// known to be 0 or 1 would be turned into a compare and a select per value.)
__device__ __forceinline__ bf16x4 relu_grad_bf16(f32x4 a, bf16x4 x) {
  unsigned one = 0x00010001u;
  asm("" : "+v"(one));
  const unsigned* xb = reinterpret_cast<const unsigned*>(&x);
  unsigned out[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const u16x2 k = __builtin_elementwise_min(__builtin_bit_cast(u16x2, xb[h]), __builtin_bit_cast(u16x2, one));
    out[h] = __builtin_bit_cast(unsigned, __builtin_bit_cast(u16x2, cvt_pk_bf16(a[2 * h], a[2 * h + 1])) * k);
  }
  return __builtin_bit_cast(bf16x4, out);
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
      // ReLU(previous) -> LDS tile [query][feature] (a lane's 4 outputs are
      // adjacent features of its query: one store)
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) *(bf16x4*)&sact[wave][r][16 * ot + 4 * g] = relu_bf16(acc[ot]);
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
// Two launches per step: the fused forward / loss / backward / weight
// gradients (k_train_fused), then the partials' deterministic reduction
// fused with the Adam update (k_train_update).
static_assert(param_offset(1) % 64 == 0 && param_offset(2) % 64 == 0 && param_offset(5) % 64 == 0 &&
                  kParams % 64 == 0,
              "parameter blocks of 64 stay within one layer");

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

// ------------------------------------------------- fused training step
// One launch computes a step's loss, its backward pass and its weight
// gradients: a block of 4 waves walks chunks of 64 samples (the batch's
// chunks c = blockIdx, blockIdx + P, ...) and keeps everything a chunk needs
// in LDS -- every layer's input and the current output gradient
// (double-buffered), all sample-major ([sample][feature]) -- so no activation
// leaves the CU. The forward and backward products read 8 consecutive
// features of a sample (one 16-byte LDS read); the weight gradients, whose
// MFMA K dimension is the samples, read 8 consecutive samples of a feature
// with two transposing LDS reads (ds_read_b64_tr_b16) from the same buffers.
// Each wave accumulates 25 of the 100 16x16 weight-gradient tiles in
// registers over all of the block's chunks and writes them once, in the
// MFMA's own lane order, as the block's partial (k_train_update sums the P
// partials in a fixed order and takes the Adam step). Layer 0's forward and
// the loss run one sample tile per wave; the hidden layers' products run a
// 2 x 2 block of tiles per wave (two 16-feature tiles of two 16-sample
// tiles), so each operand read from LDS or L2 feeds two products.
constexpr int kFChunk = 64;            // samples per chunk
constexpr int kFStride0 = 128 + 16;    // encoding rows: 128 features + pad (bf16)
constexpr int kFStride = 64 + 8;       // hidden rows: 64 features + pad (bf16)
constexpr int kFTiles = 25;            // weight-gradient tiles per wave
constexpr int kTileBase[kLayers + 1] = {0, 8, 12, 16, 20, 24, 25};  // a wave's tiles of layer l

struct FusedLds {
  bf16x8 w0[16 * 64];                  // layer 0's forward weight fragments (16 per lane group)
  __bf16 x0[kFChunk * kFStride0];      // X_0 (encoding)
  __bf16 xs[5][kFChunk * kFStride];    // X_1..X_5 (inputs of layers 1..5)
  __bf16 zs[2][kFChunk * kFStride];    // current / next output gradient
};
static_assert(sizeof(FusedLds) <= 160 * 1024, "fused training LDS exceeds the CU's 160 KB");

__device__ __forceinline__ bf16x8 lds8(const __bf16* p) { return *(const bf16x8*)p; }

typedef __attribute__((address_space(3))) bf16x4 lds_bf16x4;

// MFMA operand with K along the rows of a sample-major LDS tile: lane (g, r)
// receives buf[s0 + 8g + j][f0 + r], j = 0..7. Each ds_read_b64_tr_b16 has
// lane (g, i) of a 16-lane group read 4 adjacent features of row
// s0 + 8g + i/4 (+4 for the second read) at column f0 + 4(i%4), and hands lane
// (g, r) column f0 + r of those 4 rows.
__device__ __forceinline__ bf16x8 lds_tr8(const __bf16* buf, int stride, int s0, int f0, int g, int r) {
  const __bf16* p = buf + (s0 + 8 * g + (r >> 2)) * stride + f0 + 4 * (r & 3);
  const bf16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)p);
  const bf16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(p + 4 * stride));
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__global__ __launch_bounds__(256, 1) void k_train_fused(const bf16x8* __restrict__ wf, const bf16x8* __restrict__ wb,
                                                        const float* __restrict__ feats,
                                                        const float* __restrict__ targets, BatchN bn, NrcBounds b,
                                                        __bf16* __restrict__ partial, float* __restrict__ loss_part,
                                                        uint32_t* step) {
  __shared__ __attribute__((aligned(16))) FusedLds L;
  // the loss reduction's 4 floats sit in X_0's row padding (columns
  // 128..135 of row 0, never read by the products); used after the last chunk
  float* const red = reinterpret_cast<float*>(&L.x0[128]);
  const uint32_t n_real = batch_size(bn);
  if (step && n_real && blockIdx.x == 0 && threadIdx.x == 0) *step += 1u;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  // this wave's 2 x 2 block of 16 x 16 tiles in the hidden layers' products
  const int ot0 = 2 * (w & 1), st0 = 2 * (w >> 1);
  const uint32_t nch = (n_real + kFChunk - 1) / kFChunk;
  const float inv_3n = 1.0f / (3.0f * (float)n_real);  // the loss is the mean over 3 n values
  f32x4 dw[kFTiles];
#pragma unroll
  for (int k = 0; k < kFTiles; ++k) dw[k] = f32x4{0, 0, 0, 0};
  float loss = 0.0f;
  // layer 0's fragments are read by every wave for every chunk: from LDS,
  // which leaves the registers they would hold to the products' operands
  stage_frags<16 * 64, 256>(L.w0, wf + fwd_frag_offset(0) * 64);
  // this lane's sample of a chunk (16 w + r; padding repeats a sample, whose
  // gradient is zero) and its targets (lanes of group 0 use channels 0..2):
  // fetched one chunk ahead, so their memory latency overlaps a chunk's work
  // (loads without branches: a branch around them would make the compiler
  // wait for them where it rejoins; past the last chunk the last sample is
  // fetched again and not used)
  auto fetch = [&](uint32_t c, RawQuery& rq, float (&tg)[3]) {
    const uint32_t q = c * kFChunk + 16 * w + r;
    const uint32_t row = q < n_real ? q : n_real - 1;
    rq = fetch_query(feats + 13 * (size_t)row);
#pragma unroll
    for (int i = 0; i < 3; ++i) tg[i] = targets[3 * (size_t)row + i];
  };
  RawQuery rq_next;
  float tg_next[3];
  if (n_real) fetch(blockIdx.x, rq_next, tg_next);
  __syncthreads();
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const uint32_t s0 = c * kFChunk;
    const RawQuery rq = rq_next;
    float tg[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) tg[i] = tg_next[i];
    fetch(c + gridDim.x, rq_next, tg_next);
    // this wave's weight fragments for the next product phase are loaded one
    // phase ahead, so their L2 latency overlaps the current phase. The
    // fragment pointers go through an empty asm statement once per chunk,
    // which keeps the loads where they are written (otherwise they are
    // hoisted out of the loop into registers the products need, or sunk to
    // their use).
    typedef const __attribute__((address_space(1))) bf16x8* gptr;
    gptr wfc = (gptr)wf;
    gptr wbc = (gptr)wb;
    asm("" : "+s"(wfc), "+s"(wbc));
    auto fwd_frags = [&](int l, bf16x8(&a)[4]) {  // a[o * 2 + kb]: output tile ot0 + o, K block kb
#pragma unroll
      for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) a[o * 2 + kb] = wfc[(fwd_frag_offset(l) + (ot0 + o) * 2 + kb) * 64 + lane];
    };
    auto bwd_frags = [&](int l, bf16x8(&a)[4]) {  // a[i * nkb + kb]: input tile ot0 + i, K block kb
      const int nkb = (kOut[l] + 31) / 32;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          if (kb < nkb) a[i * nkb + kb] = wbc[(bwd_frag_offset(l) + (ot0 + i) * nkb + kb) * 64 + lane];
    };
    bf16x8 wn[4];
    fwd_frags(1, wn);
    // encode + layer 0 (this wave's sample tile w): Z_0^T = W_0^T X_0^T
    {
      const Query qq = query_of(rq, b);
      const int sl = 16 * w + r;
      f32x4 acc[4];
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) acc[ot] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        // the 4 weight fragments first: their LDS latency overlaps the encoding
        bf16x8 wa[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) wa[ot] = L.w0[(ot * 4 + kb) * 64 + lane];
        __builtin_amdgcn_sched_barrier(0);
        const bf16x8 bx = encode_frag(qq, kb, g);
        *(bf16x8*)&L.x0[sl * kFStride0 + 32 * kb + 8 * g] = bx;
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc[ot] = mfma(wa[ot], bx, acc[ot]);
      }
      // C[4g + i][r] = Z_0^T[16 ot + 4g + i][16 w + r] -> X_1 = ReLU (the
      // lane's 4 features are adjacent in the sample's row: one store)
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) *(bf16x4*)&L.xs[0][sl * kFStride + 16 * ot + 4 * g] = relu_bf16(acc[ot]);
    }
    __syncthreads();
    // layers 1..4: this wave's 2 x 2 block of tiles -- output tiles ot0,
    // ot0 + 1 of sample tiles st0, st0 + 1 -- so each operand read from LDS
    // feeds two products
#pragma unroll
    for (int l = 1; l < kLayers - 1; ++l) {
      f32x4 acc[2][2];
#pragma unroll
      for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[o][s] = f32x4{0, 0, 0, 0};
      // all 4 operand tiles are read before the first product, so the
      // products wait for one LDS round trip, not four
      bf16x8 bx[2][2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int s = 0; s < 2; ++s)
          bx[kb][s] = lds8(&L.xs[l - 1][(16 * (st0 + s) + r) * kFStride + 32 * kb + 8 * g]);
      bf16x8 wc[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) wc[k] = wn[k];
      if (l + 1 < kLayers - 1) {
        fwd_frags(l + 1, wn);
      } else {  // layer 5's two fragments (one 16-output tile)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) wn[kb] = wfc[(fwd_frag_offset(5) + kb) * 64 + lane];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int o = 0; o < 2; ++o)
#pragma unroll
          for (int s = 0; s < 2; ++s) acc[o][s] = mfma(wc[o * 2 + kb], bx[kb][s], acc[o][s]);
#pragma unroll
      for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s)
          *(bf16x4*)&L.xs[l][(16 * (st0 + s) + r) * kFStride + 16 * (ot0 + o) + 4 * g] = relu_bf16(acc[o][s]);
      __syncthreads();
    }
    // layer 5 (16 outputs, RGB used) for sample tile w, and the relative-L2
    // loss gradient (p^2 held constant): dZ_5, zero past RGB and for padding
    {
      f32x4 p = {0, 0, 0, 0};
      bf16x8 x5[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) x5[kb] = lds8(&L.xs[4][(16 * w + r) * kFStride + 32 * kb + 8 * g]);
      const bf16x8 w5a = wn[0], w5b = wn[1];
      bwd_frags(5, wn);
      p = mfma(w5a, x5[0], p);
      p = mfma(w5b, x5[1], p);
      const int sl = 16 * w + r;
      const uint32_t q = s0 + sl;
      bf16x4 v;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ch = 4 * g + i;
        float gval = 0.0f;
        if (ch < 3 && q < n_real) {
          const float pv = p[i], t = tg[i];  // ch == i on the lanes of group 0
          const float inv = 1.0f / (pv * pv + 0.01f);  // one division per value
          const float d = pv - t;
          loss += d * d * inv;
          gval = 2.0f * d * inv * inv_3n;
        }
        v[i] = to_bf16(gval);
      }
      *(bf16x4*)&L.zs[0][sl * kFStride + 4 * g] = v;
    }
    __syncthreads();
    // backward, layers 5..0: dW_l += X_l^T dZ_l over the chunk, then
    // dZ_{l-1} = (W_l dZ_l^T) * (X_l > 0)
#pragma unroll
    for (int l = kLayers - 1; l >= 0; --l) {
      const int cur = (kLayers - 1 - l) & 1;
      const __bf16* xl = l == 0 ? L.x0 : L.xs[l - 1];
      const int xstride = l == 0 ? kFStride0 : kFStride;
      // this wave's weight-gradient tiles of layer l (output tile ot_o, input
      // tile it_t, held at dw[kTileBase[l] + o * nI + t]): layer 0 output
      // tiles ot0, ot0 + 1 by input tiles 4 (w / 2) .. + 3; layers 1..4 output
      // tiles ot0, ot0 + 1 by input tiles st0, st0 + 1; layer 5 input tile w
      const int nO = l == kLayers - 1 ? 1 : 2, nI = l == 0 ? 4 : (l == kLayers - 1 ? 1 : 2);
      bf16x8 wc[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) wc[k] = wn[k];
      if (l >= 2) bwd_frags(l - 1, wn);
      // weight gradients: dZ tiles (B) and input tiles (A), K = the chunk's
      // samples, all read before the first product
      bf16x8 zb[2][2], xa[4][2];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        if (o >= nO) break;
        const int ot = l == kLayers - 1 ? 0 : ot0 + o;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) zb[o][kb] = lds_tr8(L.zs[cur], kFStride, 32 * kb, 16 * ot, g, r);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= nI) break;
        const int it = l == 0 ? 4 * (w >> 1) + t : (l == kLayers - 1 ? w : st0 + t);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) xa[t][kb] = lds_tr8(xl, xstride, 32 * kb, 16 * it, g, r);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        if (o >= nO) break;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (t >= nI) break;
          f32x4& d = dw[kTileBase[l] + o * nI + t];
#pragma unroll
          for (int kb = 0; kb < 2; ++kb) d = mfma(xa[t][kb], zb[o][kb], d);
        }
      }
      if (l >= 1) {
        // dX_l^T = W_l dZ_l^T for this wave's 2 x 2 block (input tiles ot0,
        // ot0 + 1 of sample tiles st0, st0 + 1), masked by X_l > 0: the dZ
        // tiles (B) and the ReLU masks (X_l[16 s + r][16 i + 4g .. + 3], one
        // load each) are read before the first product
        const int nkb = (kOut[l] + 31) / 32;
        bf16x4 mask[2][2];
        bf16x8 bz[2][2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            if (kb >= nkb) break;
            if (kOut[l] >= 32 * kb + 8 * g + 8) {
              bz[kb][s] = lds8(&L.zs[cur][(16 * (st0 + s) + r) * kFStride + 32 * kb + 8 * g]);
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) bz[kb][s][j] = to_bf16(0.0f);  // K past the layer's 16 outputs
            }
          }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int s = 0; s < 2; ++s)
            mask[i][s] = *(const bf16x4*)&xl[(16 * (st0 + s) + r) * kFStride + 16 * (ot0 + i) + 4 * g];
        __builtin_amdgcn_sched_barrier(0);
        f32x4 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int s = 0; s < 2; ++s) acc[i][s] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
          if (kb >= nkb) break;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int s = 0; s < 2; ++s) acc[i][s] = mfma(wc[i * nkb + kb], bz[kb][s], acc[i][s]);
        }
        // C[4g + j][r] = dX_l^T[16 i + 4g + j][16 s + r]
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int s = 0; s < 2; ++s)
            *(bf16x4*)&L.zs[cur ^ 1][(16 * (st0 + s) + r) * kFStride + 16 * (ot0 + i) + 4 * g] =
                relu_grad_bf16(acc[i][s], mask[i][s]);
      }
      __syncthreads();
    }
  }
  // this block's partial: tile k of wave w is tile 25 w + k, stored as the
  // MFMA holds it (lane-major, 4 values a lane: one 512-byte store per tile)
  // and rounded to bf16 like the products' operands (half the bytes of the
  // partials' round trip through memory, the largest fixed cost of a step)
  bf16x4* dst = reinterpret_cast<bf16x4*>(partial + (size_t)blockIdx.x * kParams) + (size_t)w * kFTiles * 64 + lane;
#pragma unroll
  for (int k = 0; k < kFTiles; ++k) {
    const unsigned pk[2] = {cvt_pk_bf16(dw[k][0], dw[k][1]), cvt_pk_bf16(dw[k][2], dw[k][3])};
    dst[k * 64] = __builtin_bit_cast(bf16x4, pk);
  }
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  if (lane == 0) red[w] = loss;
  __syncthreads();
  if (threadIdx.x == 0) loss_part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Parameter index (W_l[ji][o] at param_offset(l) + ji * kOut[l] + o) of
// position q of a partial: tile t = q / 256 (wave t / 25, its tile k = t % 25),
// lane (g, r) = (q / 4) % 64, accumulator element i = q % 4, which holds
// dW_l[16 it + 4g + i][16 ot + r] for the wave's tile assignment in
// k_train_fused.
__device__ __forceinline__ int partial_param(int q, int* layer) {
  const int t = q >> 8, lane = (q >> 2) & 63, i = q & 3;
  const int wv = t / kFTiles, k = t % kFTiles;
  int l = 0;
  while (l + 1 < kLayers && k >= kTileBase[l + 1]) ++l;
  const int nI = l == 0 ? 4 : (l == kLayers - 1 ? 1 : 2);
  const int o = (k - kTileBase[l]) / nI, tt = (k - kTileBase[l]) % nI;
  const int ot = l == kLayers - 1 ? 0 : 2 * (wv & 1) + o;
  const int it = l == 0 ? 4 * (wv >> 1) + tt : (l == kLayers - 1 ? wv : 2 * (wv >> 1) + tt);
  *layer = l;
  return param_offset(l) + (16 * it + 4 * (lane >> 4) + i) * kOut[l] + 16 * ot + (lane & 15);
}

// grad at partial position q (block of 64 positions, one tile) = the FP32 sum
// of the P bf16 partials: wave w of 16 adds partials w, w + 16, ... (16 loads
// in flight, summed pairwise, then the groups in order) and the 16 wave sums
// are added in wave order -- a fixed order, so steps are reproducible.
constexpr int kRedWaves = 16;

__device__ __forceinline__ float sum_partials(const __bf16* p, uint32_t P) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ float red[kRedWaves][64];
  float s = 0.0f;
  uint32_t c = wave;
  for (; c + 15 * kRedWaves < P; c += 16 * kRedWaves) {
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = from_bf16(p[(size_t)(c + kRedWaves * k) * kParams]);
#pragma unroll
    for (int h = 8; h > 0; h >>= 1)
#pragma unroll
      for (int k = 0; k < h; ++k) v[k] = v[2 * k] + v[2 * k + 1];
    s += v[0];
  }
  for (; c < P; c += kRedWaves) s += from_bf16(p[(size_t)c * kParams]);
  red[wave][lane] = s;
  __syncthreads();
  float t = 0.0f;
  if (wave == 0) {
    t = red[0][lane];
#pragma unroll
    for (int k = 1; k < kRedWaves; ++k) t += red[k][lane];
  }
  return t;  // valid on wave 0
}

__global__ __launch_bounds__(64 * kRedWaves) void k_train_reduce(const __bf16* partial, uint32_t P, float* grad) {
  const int q = blockIdx.x * 64 + (threadIdx.x & 63);
  const float s = sum_partials(partial + q, P);
  int l;
  if (threadIdx.x < 64) grad[partial_param(q, &l)] = s;
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
// number t is *step (counted by k_train_fused); an empty batch changes nothing.
// Block 0 also adds the batch's mean loss to *loss_dev and counts the step
// in *steps_dev when they are given.
__global__ __launch_bounds__(64 * kRedWaves) void k_train_update(const __bf16* partial, uint32_t P, float* grad,
                                                                 float* w, float* m, float* v, AdamArgs a,
                                                                 const uint32_t* step, BatchN bn, bf16x8* wf,
                                                                 bf16x8* wb, const float* loss_part, uint32_t nlp,
                                                                 float* loss_dev, uint32_t* steps_dev) {
  const uint32_t n = batch_size(bn);
  if (!n) return;
  const int q = blockIdx.x * 64 + (threadIdx.x & 63);
  const float gr = sum_partials(partial + q, P);
  if (threadIdx.x < 64) {
    const float t = (float)*step;
    a.c1 = 1.0f - powf(a.b1, t);
    a.c2 = 1.0f - powf(a.b2, t);
    int l;
    const int i = partial_param(q, &l);
    grad[i] = gr;
    float mm = m[i], vv = v[i];
    const float wn = adam_update(mm, vv, w[i], gr, a);
    m[i] = mm;
    v[i] = vv;
    w[i] = wn;
    const __bf16 hb = to_bf16(wn);
    const int in = kIn[l], out = kOut[l], j = i - param_offset(l);
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
    if (threadIdx.x < 256)
      for (uint32_t k = threadIdx.x; k < nlp; k += 256) t += loss_part[k];
    if (threadIdx.x < 256) lred[threadIdx.x] = t;
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

// Blocks of k_train_fused for a batch of at most n samples: one per CU
// (its LDS fills one), no more than the batch's 64-sample chunks. Each
// writes one weight-gradient partial (50 KB).
inline unsigned fused_grid(int cus, uint32_t n) {
  return std::max(1u, std::min(blocks(n, kFChunk), (unsigned)std::max(1, cus)));
}

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
  // training scratch: weight-gradient partials and loss sums (one per CU),
  // and the host entry points' staging of feats / targets (the largest batch)
  uint32_t cap = 0;
  __bf16* part = nullptr;
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
  dev_free(N->feats);
  dev_free(N->targets);
  N->cap = need;
  N->feats = dev_alloc<float>((size_t)need * 13);
  N->targets = dev_alloc<float>((size_t)need * 3);
}

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
// most bn.n samples (feats/targets are device arrays) in one launch; step
// counts the Adam step on the device. Returns the number of partials (and
// of per-block loss sums in N->loss_part).
static unsigned train_pass(pyre_nrc* N, const float* feats, const float* targets, const BatchN& bn, uint32_t* step,
                           hipStream_t st) {
  require(bn.n > 0, "empty training batch");
  const unsigned grid = fused_grid(N->ctx->cus, bn.n);
  hipLaunchKernelGGL(k_train_fused, dim3(grid), dim3(256), 0, st, N->wf, N->wb, feats, targets, bn, N->bounds,
                     N->part, N->loss_part, step);
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
  hipLaunchKernelGGL(k_train_reduce, dim3(kParams / 64), dim3(64 * kRedWaves), 0, st, N->part, nlp, N->grad);
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
  hipLaunchKernelGGL(k_train_update, dim3(kParams / 64), dim3(64 * kRedWaves), 0, st, N->part, nlp, N->grad, N->w,
                     N->m, N->v, adam_args(N), (const uint32_t*)N->d_step, bn, N->wf, N->wb, N->loss_part, nlp,
                     loss_dev, (uint32_t*)nullptr);
  return loss_dev || !read_back ? 0.0f : read_loss(N, nlp, n, st);
}

void nrc_train_step_counted(pyre_nrc* N, const float* feats, const float* targets, const uint32_t* count,
                            uint32_t cap, uint32_t off, uint32_t batch, hipStream_t st, float* loss_dev,
                            uint32_t* steps_dev) {
  const BatchN bn{count, cap, off, batch};
  const unsigned nlp = train_pass(N, feats + 13 * (size_t)off, targets + 3 * (size_t)off, bn, N->d_step, st);
  hipLaunchKernelGGL(k_train_update, dim3(kParams / 64), dim3(64 * kRedWaves), 0, st, N->part, nlp, N->grad, N->w,
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
    // weight-gradient partials and loss sums of k_train_fused: one per CU
    N->part = dev_alloc<__bf16>((size_t)std::max(1, ctx->cus) * kParams);
    N->loss_part = dev_alloc<float>((size_t)std::max(1, ctx->cus));
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
