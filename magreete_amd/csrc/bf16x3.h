// BF16x3 building blocks: a float x = x_hi + x_lo with x_hi = bf16(x) and
// x_lo = bf16(x - x_hi) (round to nearest even), and the complex
// multiply-accumulate of 16 x 16 tiles from twelve bf16 MFMA products
// (hi*hi + hi*lo + lo*hi for each of the four real products), accumulated in
// FP32 by v_mfma_f32_16x16x32_bf16.  About 16 significant bits survive per
// product (relative error ~1e-5).
// MFMA lane layout (measured on gfx950): lane l feeds A[l%16][8(l/16)+e] and
// B[8(l/16)+e][l%16], e = 0..7, and item r of its result is C[4(l/16)+r][l%16].
#pragma once
#include <hip/hip_runtime.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline unsigned short mg_bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// hi/lo split of one float into bf16 bit patterns
__device__ inline void mg_split(float x, unsigned short& hi, unsigned short& lo) {
  hi = mg_bf16_rne(x);
  const float xh = __uint_as_float((unsigned)hi << 16);
  lo = mg_bf16_rne(x - xh);
}

// eight bf16 patterns packed as one 16-byte vector
__device__ inline u32x4 mg_pack8(const unsigned short* p) {
  u32x4 v;
  v.x = p[0] | ((unsigned)p[1] << 16);
  v.y = p[2] | ((unsigned)p[3] << 16);
  v.z = p[4] | ((unsigned)p[5] << 16);
  v.w = p[6] | ((unsigned)p[7] << 16);
  return v;
}

__device__ inline bf16x8 mg_ld8(const unsigned short* p) {
  u32x4 v = *reinterpret_cast<const u32x4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__device__ inline bf16x8 mg_as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// sign flip of eight bf16 values (exact)
__device__ inline bf16x8 mg_neg8(bf16x8 x) {
  u32x4 v = __builtin_bit_cast(u32x4, x);
  v ^= 0x80008000u;
  return __builtin_bit_cast(bf16x8, v);
}

__device__ inline v4f mg_mfma_bf16(bf16x8 a, bf16x8 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// (r, i) += (ar + i ai) (br + i bi) for one 16 x 16 x 32 complex tile step,
// every operand given as (hi, lo) bf16 planes.
__device__ inline void mg_cmac12(bf16x8 arh, bf16x8 arl, bf16x8 aih, bf16x8 ail, bf16x8 brh, bf16x8 brl,
                                 bf16x8 bih, bf16x8 bil, v4f& r, v4f& i) {
  const bf16x8 naih = mg_neg8(aih), nail = mg_neg8(ail);
  r = mg_mfma_bf16(arh, brh, r);
  r = mg_mfma_bf16(arh, brl, r);
  r = mg_mfma_bf16(arl, brh, r);
  r = mg_mfma_bf16(naih, bih, r);
  r = mg_mfma_bf16(naih, bil, r);
  r = mg_mfma_bf16(nail, bih, r);
  i = mg_mfma_bf16(arh, bih, i);
  i = mg_mfma_bf16(arh, bil, i);
  i = mg_mfma_bf16(arl, bih, i);
  i = mg_mfma_bf16(aih, brh, i);
  i = mg_mfma_bf16(aih, brl, i);
  i = mg_mfma_bf16(ail, brh, i);
}
