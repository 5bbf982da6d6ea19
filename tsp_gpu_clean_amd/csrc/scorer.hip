// Learned candidate-edge scorer (gfx950): feature builder, a 12 -> 128 -> 128 -> 1 MLP with bf16
// weights and fp32 accumulation on MFMA (forward and backward), Adam, and top-k candidate selection.
//
// Parameters live in one fp32 buffer (master copy) and one bf16 buffer with the same layout:
//   W1 [128][16] (input features 12..15 are zero padding), b1 [128], W2 [128][128], b2 [128],
//   w3 [128], b3 [1]   -> kScorerParams floats.
// bf16 is round-to-nearest-even of the fp32 value.
//
// MFMA: v_mfma_f32_32x32x16_bf16, A 32x16, B 16x32, D 32x32 fp32.  Lane l (s = l % 32, h = l / 32)
// holds A[row s][k 8h..8h+7], B[k 8h..8h+7][col s] and D[row 8(i/4) + 4h + (i%4)][col s] for
// i = 0..15.  Activations are kept transposed (units x samples), so a layer's D registers are
// directly the next layer's B operand once the K index is permuted the same way in A: K slot q of
// lane half h in chunk (T, c) is unit 32T + 16c + 8(q/4) + 4h + (q%4).
#include "common.hpp"

namespace tspmi {

constexpr int kH = 128;            // hidden width
constexpr int kF = 16;             // padded feature width (12 used)
constexpr int kFeat = 12;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kF;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kScorerParams = kOffB3 + 1;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// fp32 -> bf16, round to nearest even (gfx950 v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)lo) | ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)hi) << 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Unit (row) of accumulator element i of lane half h in output tile T.
__device__ __forceinline__ int dunit(int T, int h, int i) { return 32 * T + 8 * (i >> 2) + 4 * h + (i & 3); }

// 8 bf16 = {W[row][c0 .. c0+3], W[row][c0+8 .. c0+11]} from a row-major bf16 matrix in LDS.
__device__ __forceinline__ bf16x8 load_perm(const uint16_t* W, int ld, int row, int c0) {
  const uint2 lo = *reinterpret_cast<const uint2*>(W + row * ld + c0);
  const uint2 hi = *reinterpret_cast<const uint2*>(W + row * ld + c0 + 8);
  return as_bf16x8(make_uint4(lo.x, lo.y, hi.x, hi.y));
}

// Pack the accumulator slots 8c..8c+7 (converted to bf16) into one B operand.
__device__ __forceinline__ bf16x8 pack_slots(const float* v) {
  return as_bf16x8(make_uint4(pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3]), pack_bf16x2(v[4], v[5]),
                              pack_bf16x2(v[6], v[7])));
}

// Row stride of W2 and W2^T in LDS, in bf16: 8 past the row, so that the 32 rows a half-wave reads
// for one MFMA operand start 16 bytes apart in the banks instead of all in one (at a stride of
// 128 elements every lane of the half-wave hit the same bank: 90 % of the forward kernel's LDS
// cycles were bank conflicts).
constexpr int kW2Ld = kH + 8;

// Weights of the forward pass in LDS (bf16 matrices, fp32 biases).
struct alignas(16) FwdLds {
  uint16_t w1[kH * kF];
  uint16_t w2[kH * kW2Ld];
  float b1[kH], b2[kH], w3[kH];
  float b3;
};

// Forward weights plus W2 transposed for the backward pass.
struct alignas(16) ScorerLds : FwdLds {
  uint16_t w2t[kH * kW2Ld];
};

// The weights into LDS with 16-byte loads, every load of a thread issued before it stores any:
// W1 (256 x 16 B), W2 (2048 x 16 B) and, for the backward pass, W2^T from its copy in global
// memory (w2t_g, written by w2t_kernel), so that no workgroup transposes in LDS.  wbf and w2t_g are
// 16-byte aligned (the entry points check it; the W1 and W2 offsets are multiples of 8 elements).
constexpr int kMlpBlock = 256;  // 4 waves, 32 samples each per step

__device__ void load_fwd(FwdLds& L, const uint16_t* __restrict__ wbf, uint16_t* w2t, const uint16_t* __restrict__ w2t_g) {
  constexpr int V1 = kH * kF / 8, V2 = kH * kH / 8, E2 = V2 / kMlpBlock;
  static_assert(V1 == kMlpBlock && V2 % kMlpBlock == 0, "weight staging assumes 256 threads");
  const int t = threadIdx.x;
  const uint4 r1 = reinterpret_cast<const uint4*>(wbf + kOffW1)[t];
  uint4 r2[E2], rt[E2];
#pragma unroll
  for (int e = 0; e < E2; ++e) r2[e] = reinterpret_cast<const uint4*>(wbf + kOffW2)[t + e * kMlpBlock];
  if (w2t != nullptr) {
#pragma unroll
    for (int e = 0; e < E2; ++e) rt[e] = reinterpret_cast<const uint4*>(w2t_g)[t + e * kMlpBlock];
  }
  float b1 = 0.0f, b2 = 0.0f, w3 = 0.0f, b3 = 0.0f;
  if (t < kH) {
    b1 = bf2f(wbf[kOffB1 + t]);
    b2 = bf2f(wbf[kOffB2 + t]);
    w3 = bf2f(wbf[kOffW3 + t]);
  }
  if (t == 0) b3 = bf2f(wbf[kOffB3]);
  reinterpret_cast<uint4*>(L.w1)[t] = r1;
  // 16 uint4 per 128-element row; row r starts at r * kW2Ld elements (16-byte aligned)
#pragma unroll
  for (int e = 0; e < E2; ++e) {
    const int v = t + e * kMlpBlock;
    reinterpret_cast<uint4*>(L.w2 + (v >> 4) * kW2Ld)[v & 15] = r2[e];
  }
  if (w2t != nullptr) {
#pragma unroll
    for (int e = 0; e < E2; ++e) {
      const int v = t + e * kMlpBlock;
      reinterpret_cast<uint4*>(w2t + (v >> 4) * kW2Ld)[v & 15] = rt[e];
    }
  }
  if (t < kH) {
    L.b1[t] = b1;
    L.b2[t] = b2;
    L.w3[t] = w3;
  }
  if (t == 0) L.b3 = b3;
  __syncthreads();
}

__device__ void load_scorer(ScorerLds& L, const uint16_t* __restrict__ wbf, const uint16_t* __restrict__ w2t_g) {
  load_fwd(L, wbf, L.w2t, w2t_g);
}

// out = W2^T (bf16 [128][128]) for the backward pass: out[c][r] = W2[r][c].
__global__ void w2t_kernel(const uint16_t* __restrict__ wbf, uint16_t* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < kH * kH) out[(x % kH) * kH + x / kH] = wbf[kOffW2 + x];
}

// Forward of 32 samples (rows base .. base+31 of X, bf16 [S][16]) by one wave.
// Returns z for sample s (valid in every lane; both halves agree); fills pre-activations.
struct Fwd {
  float h1[4][16];  // layer-1 pre-activation (units x sample s)
  float h2[4][16];  // layer-2 pre-activation
  float z;
};

__device__ __forceinline__ void forward32(const FwdLds& L, const uint16_t* __restrict__ X, int base, int S, Fwd& f) {
  const int lane = threadIdx.x & 63;
  const int s = lane & 31, h = lane >> 5;
  const int row = base + s < S ? base + s : S - 1;
  const bf16x8 xb = as_bf16x8(*reinterpret_cast<const uint4*>(X + (size_t)row * kF + 8 * h));
  // layer 1: H1^T = W1 . X^T
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    const bf16x8 a = as_bf16x8(*reinterpret_cast<const uint4*>(L.w1 + (32 * T + s) * kF + 8 * h));
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xb, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) f.h1[T][i] = acc[i] + L.b1[dunit(T, h, i)];
  }
  // layer 2: H2^T = W2 . relu(H1)^T
  bf16x8 bop[4][2];
#pragma unroll
  for (int T = 0; T < 4; ++T) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = fmaxf(f.h1[T][8 * c + q], 0.0f);
      bop[T][c] = pack_slots(v);
    }
  }
  float part = 0.0f;
#pragma unroll
  for (int To = 0; To < 4; ++To) {
    f32x16 acc = {};
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load_perm(L.w2, kW2Ld, 32 * To + s, 32 * T + 16 * c + 4 * h), bop[T][c],
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int u = dunit(To, h, i);
      f.h2[To][i] = acc[i] + L.b2[u];
      part += L.w3[u] * fmaxf(f.h2[To][i], 0.0f);
    }
  }
  f.z = part + __shfl_xor(part, 32, 64) + L.b3;
}

// Inference forward of 32 samples by one wave, same arithmetic as forward32 but keeping only the
// packed layer-1 activations (no pre-activation arrays), so the kernel runs at high occupancy.
__device__ __forceinline__ float forward32_z(const FwdLds& L, const uint16_t* __restrict__ X, int base, int S) {
  const int lane = threadIdx.x & 63;
  const int s = lane & 31, h = lane >> 5;
  const int row = base + s < S ? base + s : S - 1;
  const bf16x8 xb = as_bf16x8(*reinterpret_cast<const uint4*>(X + (size_t)row * kF + 8 * h));
  bf16x8 bop[4][2];
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    const bf16x8 a = as_bf16x8(*reinterpret_cast<const uint4*>(L.w1 + (32 * T + s) * kF + 8 * h));
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xb, acc, 0, 0, 0);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = fmaxf(acc[8 * c + q] + L.b1[dunit(T, h, 8 * c + q)], 0.0f);
      bop[T][c] = pack_slots(v);
    }
  }
  float part = 0.0f;
#pragma unroll 1
  for (int To = 0; To < 4; ++To) {
    f32x16 acc = {};
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load_perm(L.w2, kW2Ld, 32 * To + s, 32 * T + 16 * c + 4 * h), bop[T][c],
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int u = dunit(To, h, i);
      part += L.w3[u] * fmaxf(acc[i] + L.b2[u], 0.0f);
    }
  }
  return part + __shfl_xor(part, 32, 64) + L.b3;
}

__global__ __launch_bounds__(kMlpBlock) __attribute__((amdgpu_waves_per_eu(3))) void mlp_forward_kernel(const uint16_t* __restrict__ X, int S,
                                                                const uint16_t* __restrict__ wbf, float* __restrict__ z) {
  __shared__ FwdLds L;
  load_fwd(L, wbf, nullptr, nullptr);
  const int wave = blockIdx.x * (kMlpBlock / 64) + threadIdx.x / 64;
  const int waves = gridDim.x * (kMlpBlock / 64);
  const int lane = threadIdx.x & 63;
  for (int base = wave * 32; base < S; base += waves * 32) {
    const float zz = forward32_z(L, X, base, S);
    if (lane < 32 && base + lane < S) z[base + lane] = zz;
  }
}

// Training pass 1: forward + backward to the per-sample deltas, one wave per 32 samples, each
// workgroup taking kBwdGroups groups of 128 samples (the weights staged once for all of them).
// Writes (unit-major, bf16) a1T = relu(h1), d2T = delta2, d1T = delta1 as [128][S]; dz[S] = delta3;
// per-block partial sums of dw3, db3 and the loss, accumulated over the block's groups in order.
constexpr int kBwdGroups = 2;
constexpr int kBwdSamples = kBwdGroups * (kMlpBlock / 64) * 32;  // samples per workgroup

__global__ __launch_bounds__(kMlpBlock) void mlp_backward_kernel(const uint16_t* __restrict__ X,
                                                                 const float* __restrict__ y, int S,
                                                                 const uint16_t* __restrict__ wbf,
                                                                 const uint16_t* __restrict__ w2t_g, float inv_batch,
                                                                 uint16_t* __restrict__ a1T, uint16_t* __restrict__ d2T,
                                                                 uint16_t* __restrict__ d1T, float* __restrict__ part) {
  __shared__ ScorerLds L;
  __shared__ float s_w3[kMlpBlock / 64][kH + 2];
  load_scorer(L, wbf, w2t_g);
  const int lane = threadIdx.x & 63, wid = threadIdx.x / 64;
  const int s = lane & 31, h = lane >> 5;
  float dw3[4][16];
  float db3 = 0.0f, loss = 0.0f;
#pragma unroll
  for (int T = 0; T < 4; ++T)
#pragma unroll
    for (int i = 0; i < 16; ++i) dw3[T][i] = 0.0f;
  // (kept in registers over the groups: accumulating dw3 in LDS group by group avoids the 25
  // spilled registers but measured slower, 70.5 against 64.7 us at 65,536 samples)
#pragma unroll 1
  for (int grp = 0; grp < kBwdGroups; ++grp) {
    const int base = blockIdx.x * kBwdSamples + (grp * (kMlpBlock / 64) + wid) * 32;
    if (base < S) {
      Fwd f;
      forward32(L, X, base, S, f);
      const bool valid = base + s < S;
      const float yy = valid ? y[base + s] : 0.0f;
      const float zz = f.z;
      const float p = 1.0f / (1.0f + __expf(-zz));
      const float d3 = valid ? (p - yy) * inv_batch : 0.0f;
      if (h == 0 && valid) loss += (fmaxf(zz, 0.0f) - zz * yy + log1pf(__expf(-fabsf(zz)))) * inv_batch;
      if (h == 0) db3 += d3;
      // delta2 = d3 * w3 * relu'(h2); dw3 += d3 * relu(h2)
      float d2[4][16];
#pragma unroll
      for (int To = 0; To < 4; ++To)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int u = dunit(To, h, i);
          const float a2 = fmaxf(f.h2[To][i], 0.0f);
          dw3[To][i] += d3 * a2;
          d2[To][i] = f.h2[To][i] > 0.0f ? d3 * L.w3[u] : 0.0f;
        }
      // dH1^T = W2^T . delta2 (MFMA), delta1 = dH1 * relu'(h1)
      bf16x8 bop[4][2];
#pragma unroll
      for (int To = 0; To < 4; ++To)
#pragma unroll
        for (int c = 0; c < 2; ++c) bop[To][c] = pack_slots(&d2[To][8 * c]);
#pragma unroll
      for (int Ti = 0; Ti < 4; ++Ti) {
        f32x16 acc = {};
#pragma unroll
        for (int To = 0; To < 4; ++To)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load_perm(L.w2t, kW2Ld, 32 * Ti + s, 32 * To + 16 * c + 4 * h),
                                                          bop[To][c], acc, 0, 0, 0);
        if (valid) {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int u = dunit(Ti, h, i);
            const float d1 = f.h1[Ti][i] > 0.0f ? acc[i] : 0.0f;
            d1T[(size_t)u * S + base + s] = f2bf(d1);
            a1T[(size_t)u * S + base + s] = f2bf(fmaxf(f.h1[Ti][i], 0.0f));
            d2T[(size_t)u * S + base + s] = f2bf(d2[Ti][i]);
          }
        }
      }
    }
  }
  // reduce dw3 over the 32 samples of the wave (lanes s), then over waves
#pragma unroll
  for (int T = 0; T < 4; ++T)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      float v = dw3[T][i];
#pragma unroll
      for (int m = 16; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
      if (s == 0) s_w3[wid][dunit(T, h, i)] = v;
    }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    db3 += __shfl_xor(db3, m, 64);
    loss += __shfl_xor(loss, m, 64);
  }
  if (lane == 0) {
    s_w3[wid][kH] = db3;
    s_w3[wid][kH + 1] = loss;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < kH + 2; x += kMlpBlock) {
    float v = 0.0f;
    for (int w = 0; w < kMlpBlock / 64; ++w) v += s_w3[w][x];
    part[(size_t)blockIdx.x * (kH + 2) + x] = v;
  }
}

// Training pass 2: weight gradients over a chunk of samples per block.
//   dW2[o][i] = sum_s d2T[o][s] a1T[i][s]   (MFMA, K = samples; wave w owns output row-tile w)
//   dW1[u][f] = sum_s d1T[u][s] X[s][f]      (VALU)
//   db1[u] = sum_s d1T[u][s], db2[u] = sum_s d2T[u][s]
// Writes a partial gradient (dW1, db1, dW2, db2) per block.
constexpr int kGradChunk = 256;

__global__ __launch_bounds__(kMlpBlock) void mlp_wgrad_kernel(const uint16_t* __restrict__ X, int S,
                                                              const uint16_t* __restrict__ a1T,
                                                              const uint16_t* __restrict__ d2T,
                                                              const uint16_t* __restrict__ d1T, float* __restrict__ part) {
  // the chunk's feature rows, read by every unit's dW1 thread below: staged once (8 KB) instead of
  // 128 loads of each row from the cache
  __shared__ uint4 sx[kGradChunk * kF / 8];
  const int lane = threadIdx.x & 63, wid = threadIdx.x / 64;
  const int s = lane & 31, h = lane >> 5;
  const int lo = blockIdx.x * kGradChunk;
  const int hi = min(lo + kGradChunk, S);
  float* out = part + (size_t)blockIdx.x * kOffW3;  // dW1, db1, dW2, db2 (W3/b3 come from pass 1)
  for (int v = threadIdx.x; v < (hi - lo) * (kF / 8); v += kMlpBlock)
    sx[v] = reinterpret_cast<const uint4*>(X + (size_t)lo * kF)[v];
  // dW2 row-tile To = wid, column tiles Ti = 0..3
  f32x16 acc[4] = {};
  const int o = 32 * wid + s;
  // whole 16-sample steps without bounds tests, four unrolled so their loads overlap
  const int full = lo + ((hi - lo) & ~15);
#pragma unroll 4
  for (int k0 = lo; k0 < full; k0 += 16) {
    const int kk = k0 + 8 * h;
    const uint4 av = *reinterpret_cast<const uint4*>(d2T + (size_t)o * S + kk);
    uint4 bv[4];
#pragma unroll
    for (int Ti = 0; Ti < 4; ++Ti) bv[Ti] = *reinterpret_cast<const uint4*>(a1T + (size_t)(32 * Ti + s) * S + kk);
#pragma unroll
    for (int Ti = 0; Ti < 4; ++Ti)
      acc[Ti] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(av), as_bf16x8(bv[Ti]), acc[Ti], 0, 0, 0);
  }
  for (int k0 = full; k0 < hi; k0 += 16) {  // the last partial step (8 samples), zero past the end
    const int kk = k0 + 8 * h;
    uint4 av = make_uint4(0, 0, 0, 0);
    if (kk + 8 <= hi) av = *reinterpret_cast<const uint4*>(d2T + (size_t)o * S + kk);
    else {
      uint16_t t[8];
      for (int q = 0; q < 8; ++q) t[q] = kk + q < hi ? d2T[(size_t)o * S + kk + q] : 0;
      av = make_uint4(t[0] | (t[1] << 16), t[2] | (t[3] << 16), t[4] | (t[5] << 16), t[6] | (t[7] << 16));
    }
#pragma unroll
    for (int Ti = 0; Ti < 4; ++Ti) {
      const int in = 32 * Ti + s;
      uint4 bv = make_uint4(0, 0, 0, 0);
      if (kk + 8 <= hi) bv = *reinterpret_cast<const uint4*>(a1T + (size_t)in * S + kk);
      else {
        uint16_t t[8];
        for (int q = 0; q < 8; ++q) t[q] = kk + q < hi ? a1T[(size_t)in * S + kk + q] : 0;
        bv = make_uint4(t[0] | (t[1] << 16), t[2] | (t[3] << 16), t[4] | (t[5] << 16), t[6] | (t[7] << 16));
      }
      acc[Ti] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(av), as_bf16x8(bv), acc[Ti], 0, 0, 0);
    }
  }
  // D[row = out unit][col = in unit]: row 32*wid + 8(i/4) + 4h + (i%4), col 32*Ti + s
#pragma unroll
  for (int Ti = 0; Ti < 4; ++Ti)
#pragma unroll
    for (int i = 0; i < 16; ++i) out[kOffW2 + dunit(wid, h, i) * kH + 32 * Ti + s] = acc[Ti][i];
  // dW1, db1, db2 on VALU: thread t covers unit u = t / 2 and features (t % 2) * 8 .. +7 (12 used);
  // samples are read 8 at a time (S and the chunk bounds are multiples of 8).
  const int u = threadIdx.x >> 1;
  const int f0 = (threadIdx.x & 1) * 8;
  const bool odd = (threadIdx.x & 1) != 0;
  float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float gb = 0.0f;  // db1 (even threads) or db2 (odd threads)
  __syncthreads();  // sx staged
#pragma unroll 4
  for (int k = lo; k < hi; k += 8) {
    const uint4 dv = *reinterpret_cast<const uint4*>(d1T + (size_t)u * S + k);
    const uint4 ev = *reinterpret_cast<const uint4*>((odd ? d2T : d1T) + (size_t)u * S + k);
    const uint32_t ds[4] = {dv.x, dv.y, dv.z, dv.w};
    const uint32_t es[4] = {ev.x, ev.y, ev.z, ev.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float d1 = bf2f((uint16_t)(ds[e >> 1] >> (16 * (e & 1))));
      gb += bf2f((uint16_t)(es[e >> 1] >> (16 * (e & 1))));
      const uint4 xv = sx[(k - lo + e) * (kF / 8) + (f0 >> 3)];
      const uint32_t xs[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        g[2 * q] += d1 * bf2f((uint16_t)(xs[q] & 0xffff));
        g[2 * q + 1] += d1 * bf2f((uint16_t)(xs[q] >> 16));
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) out[kOffW1 + u * kF + f0 + q] = (f0 + q < kFeat) ? g[q] : 0.0f;
  out[(odd ? kOffB2 : kOffB1) + u] = gb;
}

// grad = sum of the per-block partials, returning the loss in grad[kScorerParams].  A workgroup
// covers 64 parameters with 16 lanes each; lane group y sums partial sets y, y+16, y+32, ... (in
// that order, eight loads in flight at a time) and the sixteen group sums are added in a fixed
// order, so the result does not depend on scheduling.
constexpr int kSumParams = 64;
constexpr int kSumGroups = 16;

__global__ __launch_bounds__(kSumParams * kSumGroups) void mlp_grad_sum_kernel(const float* __restrict__ wpart, int nw,
                                                                            const float* __restrict__ p1part, int n1,
                                                                            float* __restrict__ grad) {
  __shared__ float acc[kSumGroups][kSumParams];
  const int tx = threadIdx.x % kSumParams, ty = threadIdx.x / kSumParams;
  const int x = blockIdx.x * kSumParams + tx;
  float v = 0.0f;
  if (x < kOffW3) {
    int b = ty;
    for (; b + 7 * kSumGroups < nw; b += 8 * kSumGroups) {
      float t[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) t[q] = wpart[(size_t)(b + kSumGroups * q) * kOffW3 + x];
#pragma unroll
      for (int q = 0; q < 8; ++q) v += t[q];
    }
    for (; b < nw; b += kSumGroups) v += wpart[(size_t)b * kOffW3 + x];
  } else if (x < kScorerParams + 1) {
    const int t = x - kOffW3;  // w3 [0, 128), b3 128, loss 129
    for (int b = ty; b < n1; b += kSumGroups) v += p1part[(size_t)b * (kH + 2) + t];
  }
  acc[ty][tx] = v;
  __syncthreads();
  if (ty == 0 && x < kScorerParams + 1) {
    float r = acc[0][tx];
#pragma unroll
    for (int y = 1; y < kSumGroups; ++y) r += acc[y][tx];
    grad[x] = r;
  }
}

// Adam on the fp32 master weights, then refresh the bf16 copy (padding features stay zero).
__global__ void adam_kernel(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ grad, int step, float lr, float b1, float b2, float eps, float wd,
                            uint16_t* __restrict__ wbf) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= kScorerParams) return;
  const bool pad = x < kOffB1 && (x % kF) >= kFeat;
  float g = pad ? 0.0f : grad[x];
  float th = theta[x];
  if (wd != 0.0f && (x < kOffB1 || (x >= kOffW2 && x < kOffB2) || (x >= kOffW3 && x < kOffB3))) g += wd * th;
  const float mm = b1 * m[x] + (1.0f - b1) * g;
  const float vv = b2 * v[x] + (1.0f - b2) * g * g;
  m[x] = mm;
  v[x] = vv;
  const float mh = mm / (1.0f - powf(b1, (float)step));
  const float vh = vv / (1.0f - powf(b2, (float)step));
  th -= lr * mh / (sqrtf(vh) + eps);
  if (pad) th = 0.0f;
  theta[x] = th;
  wbf[x] = f2bf(th);
}

__global__ void to_bf16_kernel(const float* __restrict__ src, int count, uint16_t* __restrict__ dst) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < count) dst[x] = f2bf(src[x]);
}

// Features of (city u, its r-th nearest neighbour v), r < 32, from coordinates, the 32-NN table
// and the distance matrix: rank, distance ratios to both endpoints' neighbourhood scales,
// reverse rank / mutuality, and the relative offset of v in units of u's 32-NN radius.
__global__ void scorer_features_kernel(const double* __restrict__ xy, int n, const int32_t* __restrict__ nn,
                                       const int32_t* __restrict__ D, int ld, uint16_t* __restrict__ X) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * 32) return;
  const int u = idx >> 5, r = idx & 31;
  const int v = clamp_id(nn[(size_t)u * 32 + r], n);
  const float d = (float)D[(size_t)u * ld + v];
  const float du1 = fmaxf((float)D[(size_t)u * ld + clamp_id(nn[(size_t)u * 32], n)], 1.0f);
  const float du8 = fmaxf((float)D[(size_t)u * ld + clamp_id(nn[(size_t)u * 32 + 7], n)], 1.0f);
  const float du32 = fmaxf((float)D[(size_t)u * ld + clamp_id(nn[(size_t)u * 32 + 31], n)], 1.0f);
  const float dv1 = fmaxf((float)D[(size_t)v * ld + clamp_id(nn[(size_t)v * 32], n)], 1.0f);
  const float dv32 = fmaxf((float)D[(size_t)v * ld + clamp_id(nn[(size_t)v * 32 + 31], n)], 1.0f);
  int rev = 32;
  for (int q = 0; q < 32; ++q)
    if (nn[(size_t)v * 32 + q] == u) {
      rev = q;
      break;
    }
  const float dx = (float)(xy[2 * v] - xy[2 * u]) / du32;
  const float dy = (float)(xy[2 * v + 1] - xy[2 * u + 1]) / du32;
  float f[16];
  f[0] = r / 31.0f;
  f[1] = d / du32;
  f[2] = d / dv32;
  f[3] = du1 / fmaxf(d, 1.0f);
  f[4] = dv1 / fmaxf(d, 1.0f);
  f[5] = rev < 32 ? rev / 31.0f : 1.5f;
  f[6] = rev < 32 ? 1.0f : 0.0f;
  f[7] = dx;
  f[8] = dy;
  f[9] = d / du8;
  f[10] = (d - du1) / fmaxf(du32 - du1, 1.0f);
  f[11] = du8 / du32;
#pragma unroll
  for (int q = 12; q < 16; ++q) f[q] = 0.0f;
  uint16_t* out = X + (size_t)idx * kF;
#pragma unroll
  for (int q = 0; q < 16; ++q) out[q] = f2bf(f[q]);
}

// Labels: 1 when (u, r-th neighbour) is an edge of the tour (either direction).
__global__ void scorer_labels_kernel(const int32_t* __restrict__ tour, int n, const int32_t* __restrict__ nn,
                                     int32_t* __restrict__ succ, float* __restrict__ y, int phase) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (phase == 0) {
    if (idx < n) succ[clamp_id(tour[idx], n)] = clamp_id(tour[idx + 1 == n ? 0 : idx + 1], n);
    return;
  }
  if (idx >= n * 32) return;
  const int u = idx >> 5;
  const int v = clamp_id(nn[idx], n);
  y[idx] = (succ[u] == v || succ[v] == u) ? 1.0f : 0.0f;
}

// Top-k of the 32 scored neighbours of every city (higher score first, lower rank on ties).
__global__ void scorer_topk_kernel(const float* __restrict__ z, const int32_t* __restrict__ nn, int n, int k,
                                   int32_t* __restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  float sc[32];
  for (int r = 0; r < 32; ++r) sc[r] = z[(size_t)u * 32 + r];
  for (int t = 0; t < k; ++t) {
    int best = -1;
    for (int r = 0; r < 32; ++r)
      if (sc[r] > -INFINITY && (best < 0 || sc[r] > sc[best])) best = r;
    out[(size_t)u * k + t] = best >= 0 ? nn[(size_t)u * 32 + best] : -1;
    if (best >= 0) sc[best] = -INFINITY;
  }
}

}  // namespace tspmi

using namespace tspmi;

extern "C" int tspmi_scorer_param_count(void) { return kScorerParams; }

extern "C" int tspmi_scorer_features(const double* xy, int n, const int32_t* nn32, const int32_t* D, int ld,
                                     uint16_t* X, void* stream) {
  TSPMI_CHECK_ARG(n >= 33 && ld >= n, "tspmi_scorer_features: need n >= 33 and ld >= n (n=%d ld=%d)", n, ld);
  TSPMI_CHECK_ARG(xy && nn32 && D && X, "tspmi_scorer_features: null pointer");
  hipLaunchKernelGGL(scorer_features_kernel, dim3((n * 32 + 255) / 256), dim3(256), 0, (hipStream_t)stream, xy, n, nn32,
                     D, ld, X);
  TSPMI_CHECK_LAUNCH("tspmi_scorer_features launch");
  return TSPMI_OK;
}

extern "C" int tspmi_scorer_labels(const int32_t* tour, int n, const int32_t* nn32, int32_t* work, float* y,
                                   void* stream) {
  TSPMI_CHECK_ARG(n >= 33 && tour && nn32 && work && y, "tspmi_scorer_labels: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(scorer_labels_kernel, dim3((n + 255) / 256), dim3(256), 0, s, tour, n, nn32, work, y, 0);
  hipLaunchKernelGGL(scorer_labels_kernel, dim3((n * 32 + 255) / 256), dim3(256), 0, s, tour, n, nn32, work, y, 1);
  TSPMI_CHECK_LAUNCH("tspmi_scorer_labels launch");
  return TSPMI_OK;
}

extern "C" int tspmi_mlp_forward(const uint16_t* X, int S, const uint16_t* wbf, float* z, void* stream) {
  TSPMI_CHECK_ARG(S >= 0, "tspmi_mlp_forward: S < 0");
  if (S == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(X && wbf && z && ((uintptr_t)X % 16) == 0 && ((uintptr_t)wbf % 16) == 0,
                  "tspmi_mlp_forward: null or misaligned pointer");
  // Persistent grid: the weights are staged into LDS once per workgroup (36 KB, four workgroups
  // per CU), and every wave then walks groups of 32 samples.
  const int waves = (S + 31) / 32;
  int blocks = (waves + 3) / 4;
  const int cap = 3 * device_cu_count();
  blocks = blocks > cap ? cap : blocks;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(kMlpBlock), 0, (hipStream_t)stream, X, S, wbf, z);
  TSPMI_CHECK_LAUNCH("tspmi_mlp_forward launch");
  return TSPMI_OK;
}

extern "C" int tspmi_mlp_grad(const uint16_t* X, const float* y, int S, const uint16_t* wbf, uint16_t* work,
                              float* partials, float* grad, void* stream) {
  TSPMI_CHECK_ARG(S > 0 && X && y && wbf && work && partials && grad && ((uintptr_t)X % 16) == 0 && (S % 8) == 0,
                  "tspmi_mlp_grad: bad arguments (S must be a positive multiple of 8)");
  hipStream_t s = (hipStream_t)stream;
  TSPMI_CHECK_ARG(((uintptr_t)wbf % 16) == 0 && ((uintptr_t)work % 16) == 0,
                  "tspmi_mlp_grad: wbf and work must be 16-byte aligned");
  uint16_t* a1T = work;
  uint16_t* d2T = work + (size_t)kH * S;
  uint16_t* d1T = work + (size_t)2 * kH * S;
  uint16_t* w2t = work + (size_t)3 * kH * S;  // 16-byte aligned: S is a multiple of 8
  const int n1 = (S + kBwdSamples - 1) / kBwdSamples;
  const int nw = (S + kGradChunk - 1) / kGradChunk;
  float* p1 = partials;
  float* pw = partials + (size_t)n1 * (kH + 2);
  hipLaunchKernelGGL(w2t_kernel, dim3(kH * kH / 256), dim3(256), 0, s, wbf, w2t);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3(n1), dim3(kMlpBlock), 0, s, X, y, S, wbf, w2t, 1.0f / (float)S, a1T, d2T,
                     d1T, p1);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(nw), dim3(kMlpBlock), 0, s, X, S, a1T, d2T, d1T, pw);
  hipLaunchKernelGGL(mlp_grad_sum_kernel, dim3((kScorerParams + 1 + kSumParams - 1) / kSumParams),
                     dim3(kSumParams * kSumGroups), 0, s, pw,
                     nw, p1, n1, grad);
  TSPMI_CHECK_LAUNCH("tspmi_mlp_grad launch");
  return TSPMI_OK;
}

extern "C" int tspmi_mlp_work_sizes(int S, long* work_elems, long* partial_floats) {
  TSPMI_CHECK_ARG(S > 0 && work_elems && partial_floats, "tspmi_mlp_work_sizes: bad arguments");
  const long n1 = (S + kBwdSamples - 1) / kBwdSamples;
  const long nw = (S + kGradChunk - 1) / kGradChunk;
  *work_elems = 3L * kH * S + (long)kH * kH;
  *partial_floats = n1 * (kH + 2) + nw * kOffW3;
  return TSPMI_OK;
}

extern "C" int tspmi_adam(float* theta, float* m, float* v, const float* grad, int step, float lr, float beta1,
                          float beta2, float eps, float weight_decay, uint16_t* wbf, void* stream) {
  TSPMI_CHECK_ARG(theta && m && v && grad && wbf && step >= 1, "tspmi_adam: bad arguments");
  hipLaunchKernelGGL(adam_kernel, dim3((kScorerParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, theta, m, v, grad,
                     step, lr, beta1, beta2, eps, weight_decay, wbf);
  TSPMI_CHECK_LAUNCH("tspmi_adam launch");
  return TSPMI_OK;
}

extern "C" int tspmi_to_bf16(const float* src, int count, uint16_t* dst, void* stream) {
  TSPMI_CHECK_ARG(count >= 0 && (count == 0 || (src && dst)), "tspmi_to_bf16: bad arguments");
  if (count == 0) return TSPMI_OK;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, src, count, dst);
  TSPMI_CHECK_LAUNCH("tspmi_to_bf16 launch");
  return TSPMI_OK;
}

extern "C" int tspmi_scorer_topk(const float* z, const int32_t* nn32, int n, int k, int32_t* out, void* stream) {
  TSPMI_CHECK_ARG(n >= 1 && k >= 1 && k <= 32 && z && nn32 && out, "tspmi_scorer_topk: bad arguments");
  hipLaunchKernelGGL(scorer_topk_kernel, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, z, nn32, n, k, out);
  TSPMI_CHECK_LAUNCH("tspmi_scorer_topk launch");
  return TSPMI_OK;
}
