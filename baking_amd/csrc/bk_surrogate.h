// Return-map surrogate network: 7 inputs -> 4 hidden layers of 64 SiLU units
// -> 4 outputs, evaluated in BF16 on the matrix cores with FP32 accumulation.
//
// Inputs (one row of 8 floats per particle, the 8th unused): the three
// principal trial log-strains of the Ding return map in descending order, the
// temperature its yield and viscosity factors are taken at, the saturation,
// the porosity before the substep and |Fp - I|_F.  Outputs: the three returned
// log-strains in the same order and the porosity increment of the substep.
//
// FP32 parameters (training): W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[64][64] b4[64] W5[4][64] b5[4], row-major [out][in], acting on
// normalised rows x_n = (x - in_shift) * in_scale and producing normalised
// outputs y_n, y = y_n * out_scale + out_shift.
//
// Inference blob (device, built by the host from the FP32 parameters):
//   float  in_shift[8], in_scale[8], out_shift[4], out_scale[4]
//   float  bias[4][64]      hidden layers 1..4
//   float  bias_out[16]     outputs 0..3, then zeros
//   uint16 w1[4][64][4]     bf16 tiles of W1, layer-1 input padded to 16
//   uint16 w2..w4[8][64][8]
//   uint16 w5[2][64][8]     W5 padded to 16 output units
// w1's tiles are 16x16 A operands of v_mfma_f32_16x16x16_bf16 in lane order:
// lane l, element i holds W1[16*mu + l%16][4*(l/16) + i] (zero outside the
// matrix), tile index mu.  The other layers use v_mfma_f32_16x16x32_bf16: tile
// (mu, kc) is a 16x32 A operand whose lane l, element j holds
// W[16*mu + l%16][unit(kc, l/16, j)] with
//   unit(kc, g, j) = 32*kc + 4*g + j          (j < 4)
//                    32*kc + 16 + 4*g + j - 4  (j >= 4),
// tile index mu*2 + kc.  The instruction pairs A's element j of lane group g
// with B's element j of lane group g (one K slot); the sum over K does not
// depend on which unit a slot stands for, so the slots follow the units a lane
// already holds (below).
//
// A layer computes H^T = W * X^T with the units as rows (A = weights) and the
// particles as columns (B = the previous layer).  The 16x16 output tile mu a
// lane holds -- D[4*(l/16) + i][l%16], units 16*mu + 4*(l/16) + i -- is the
// B operand of 16x16x16 as it is, and tiles 2*kc and 2*kc + 1 side by side
// are the B operand of chunk kc of 16x16x32 in the unit order above:
// activations never leave registers between layers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bk {

constexpr int kMlpIn = 7;
constexpr int kMlpRow = 8;   // floats per feature row
constexpr int kMlpWidth = 64;
constexpr int kMlpOut = 4;
constexpr int kMlpHidden = 4;  // hidden layers

// FP32 parameter offsets
constexpr int kThetaW1 = 0;
constexpr int kThetaB1 = kThetaW1 + kMlpWidth * kMlpRow;
constexpr int kThetaHidden = kThetaB1 + kMlpWidth;  // W2 b2 W3 b3 W4 b4
constexpr int kThetaHiddenStride = kMlpWidth * kMlpWidth + kMlpWidth;
constexpr int kThetaW5 = kThetaHidden + 3 * kThetaHiddenStride;
constexpr int kThetaB5 = kThetaW5 + kMlpOut * kMlpWidth;
constexpr int kThetaCount = kThetaB5 + kMlpOut;  // 13316

// Blob offsets in bytes
constexpr int kBlobInShift = 0;
constexpr int kBlobInScale = kBlobInShift + 32;
constexpr int kBlobOutShift = kBlobInScale + 32;
constexpr int kBlobOutScale = kBlobOutShift + 16;
constexpr int kBlobBias = kBlobOutScale + 16;              // float[4][64]
constexpr int kBlobBiasOut = kBlobBias + 4 * 64 * 4;       // float[16]
constexpr int kBlobW1 = kBlobBiasOut + 16 * 4;             // 4 tiles
constexpr int kBlobTile = 64 * 8;                          // bytes per 16x16 bf16 tile
constexpr int kBlobTile32 = 64 * 16;                       // bytes per 16x32 bf16 tile
constexpr int kBlobW2 = kBlobW1 + 4 * kBlobTile;           // 8 tiles each
constexpr int kBlobW5 = kBlobW2 + 3 * 8 * kBlobTile32;     // 2 tiles
constexpr int kBlobBytes = kBlobW5 + 2 * kBlobTile32;      // 29856

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// 16x16x32 BF16 product with FP32 accumulation; operands as bf16 bit patterns.
__device__ __forceinline__ floatx4 mfma32(const bf16x8 a, const bf16x8 b, const floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mfma_bf16x8, a), __builtin_bit_cast(mfma_bf16x8, b), c, 0, 0, 0);
}
// B operand of chunk kc: output tiles 2*kc and 2*kc + 1 of the previous layer side by side.
__device__ __forceinline__ bf16x8 chunk(const bf16x4 lo, const bf16x4 hi) {
  return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// FP32 -> BF16 bits, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ short to_bf16(float x) { return __builtin_bit_cast(short, (__bf16)x); }

// SiLU as used by training (IEEE exp and divide).
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

// SiLU for inference: hardware exp2 and reciprocal (a few ulps), its result is rounded to BF16 next.
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// Evaluate the network for the 64 particles of one wavefront.  `xin` is this
// wavefront's LDS staging area of 64 rows x 8 floats holding the normalised
// inputs (row = lane); on return rows hold the normalised outputs in
// columns 0..3.  `blob` is the inference blob (in LDS or global memory).
// Every lane of the wavefront must call it.
__device__ inline void mlp_wave(const uint8_t* blob, float* xin) {
  const int lane = int(__lane_id());
  const int col = lane & 15, grp = lane >> 4;
  const float* bias = reinterpret_cast<const float*>(blob + kBlobBias);
  const float* bias_out = reinterpret_cast<const float*>(blob + kBlobBiasOut);
  const bf16x4* w1 = reinterpret_cast<const bf16x4*>(blob + kBlobW1);
  const bf16x8* w5 = reinterpret_cast<const bf16x8*>(blob + kBlobW5);

  // layer-1 B operands: particle 16*pt + col, inputs 4*grp .. 4*grp+3 (grp >= 2: padding)
  bf16x4 h[4][4];  // [k tile][particle tile]
#pragma unroll
  for (int pt = 0; pt < 4; ++pt) {
    bf16x4 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = grp < 2 ? to_bf16(xin[(16 * pt + col) * kMlpRow + 4 * grp + i]) : short(0);
    h[0][pt] = b;
  }
  __builtin_amdgcn_wave_barrier();
  // layer 1: one k tile
  floatx4 acc[4][4];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    const bf16x4 a = w1[mu * 64 + lane];
    floatx4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = bias[16 * mu + 4 * grp + i];
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) acc[mu][pt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h[0][pt], c, 0, 0, 0);
  }
  // hidden layers 2..4
#pragma unroll
  for (int layer = 0; layer < 3; ++layer) {
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) {
        bf16x4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = to_bf16(silu_fast(acc[mu][pt][i]));
        h[mu][pt] = b;
      }
    const bf16x8* w = reinterpret_cast<const bf16x8*>(blob + kBlobW2 + layer * 8 * kBlobTile32);
    const float* bl = bias + 64 * (layer + 1);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      floatx4 c;
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = bl[16 * mu + 4 * grp + i];
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) acc[mu][pt] = c;
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) {
        const bf16x8 a = w[(mu * 2 + kc) * 64 + lane];
#pragma unroll
        for (int pt = 0; pt < 4; ++pt) acc[mu][pt] = mfma32(a, chunk(h[2 * kc][pt], h[2 * kc + 1][pt]), acc[mu][pt]);
      }
    }
  }
  // output layer: 16 units, 4 real
#pragma unroll
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) {
      bf16x4 b;
#pragma unroll
      for (int i = 0; i < 4; ++i) b[i] = to_bf16(silu_fast(acc[mu][pt][i]));
      h[mu][pt] = b;
    }
  floatx4 out[4];
#pragma unroll
  for (int pt = 0; pt < 4; ++pt) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[pt][i] = bias_out[4 * grp + i];
  }
#pragma unroll
  for (int kc = 0; kc < 2; ++kc) {
    const bf16x8 a = w5[kc * 64 + lane];
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) out[pt] = mfma32(a, chunk(h[2 * kc][pt], h[2 * kc + 1][pt]), out[pt]);
  }
  // lanes 0..15 hold output units 0..3 of particles 16*pt + col
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (grp == 0) {
#pragma unroll
    for (int pt = 0; pt < 4; ++pt)
#pragma unroll
      for (int i = 0; i < 4; ++i) xin[(16 * pt + col) * kMlpRow + i] = out[pt][i];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Network outputs for `count` feature rows (rows [first, first + count) of
// `features`), one wavefront per 64 rows; the blob is staged in LDS once per
// workgroup.  Launch with 256 threads and kBlobBytes + 4 * 64 * 8 * 4 bytes
// of dynamic LDS.  Every lane of a workgroup runs the same number of loop
// iterations.
__device__ inline void mlp_rows(const uint8_t* __restrict__ blob_g, const float* __restrict__ features, float* __restrict__ out,
                                int count, uint8_t* lds) {
  uint8_t* blob = lds;
  float* stage = reinterpret_cast<float*>(lds + kBlobBytes) + (threadIdx.x >> 6) * 64 * kMlpRow;
  for (int i = threadIdx.x; i < kBlobBytes / 16; i += blockDim.x)
    reinterpret_cast<uint4*>(blob)[i] = reinterpret_cast<const uint4*>(blob_g)[i];
  __syncthreads();
  const float* in_shift = reinterpret_cast<const float*>(blob + kBlobInShift);
  const float* in_scale = reinterpret_cast<const float*>(blob + kBlobInScale);
  const float* out_shift = reinterpret_cast<const float*>(blob + kBlobOutShift);
  const float* out_scale = reinterpret_cast<const float*>(blob + kBlobOutScale);
  const int lane = int(__lane_id());
  const int waves = int(gridDim.x) * int(blockDim.x >> 6);
  for (int base = (int(blockIdx.x) * int(blockDim.x >> 6) + int(threadIdx.x >> 6)) * 64; base < count; base += waves * 64) {
    const int r = base + lane;
#pragma unroll
    for (int k = 0; k < kMlpRow; ++k)
      stage[lane * kMlpRow + k] = (r < count && k < kMlpIn) ? (features[size_t(r) * kMlpRow + k] - in_shift[k]) * in_scale[k] : 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    mlp_wave(blob, stage);
    if (r < count) {
#pragma unroll
      for (int k = 0; k < kMlpOut; ++k) out[size_t(r) * kMlpOut + k] = stage[lane * kMlpRow + k] * out_scale[k] + out_shift[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

constexpr int kMlpLdsBytes = kBlobBytes + 4 * 64 * kMlpRow * 4;

}  // namespace bk
