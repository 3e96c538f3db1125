// Shared definitions for the coupled-dipole HIP library (gfx950).
#pragma once

#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MG_HD __host__ __device__ inline
#else
#define MG_HD inline
#endif

// Status codes returned by every entry point (0 ok, >0 LAPACK-style info,
// <0 invalid argument or runtime failure).
enum {
  MG_OK = 0,
  MG_ERR_ARG = -1,
  MG_ERR_HIP = -2,
  MG_ERR_KIND = -3,
  MG_ERR_NOMEM = -4,
  MG_ERR_SYNC = -5,      // a grid barrier did not complete within its spin limit
  MG_ERR_UNSUPPORTED = -6,
};

// Green's-function kinds, one per reference class.
enum {
  MG_2D_TM = 0,      // Transmission2D_scalar, d = 1
  MG_2D_TE = 1,      // Transmission2D_vector, d = 2
  MG_3D_VECTOR = 2,  // Transmission3D_vector, d = 3
  MG_3D_SCALAR = 3,  // Transmission3D_scalar, d = 1
};

// Element precisions of stored matrices and vectors.
enum {
  MG_C128 = 0,   // complex double, interleaved
  MG_C64 = 1,    // complex float, interleaved
  MG_BF16X3 = 2, // complex float storage, products as three bf16 MFMAs
  MG_BF16X6 = 3, // complex float storage, products as six bf16 MFMAs (three-part splits)
};

// Assembly flags.
enum {
  MG_F_REGULARIZE = 1,   // R < radius -> R = 0 (reference quirks per kind)
  MG_F_ZERO_SELF = 2,    // block (i, i) -> 0 (rows and cols are the same point set)
  MG_F_DIAG = 4,         // global diagonal entry (r, r) -> diag value
  MG_F_NANFILL = 8,      // NaN entries -> fill value (LDOS measurement fix-up)
  MG_F_PERIODIC_X = 16,
  MG_F_PERIODIC_Y = 32,
  MG_F_PERIODIC_Z = 64,
  MG_F_BESSEL_EXACT = 128,  // 2D kinds: accurate Bessel functions instead of the reference's
};

MG_HD int mg_kind_dim(int kind) {
  return kind == MG_2D_TE ? 2 : (kind == MG_3D_VECTOR ? 3 : 1);
}
MG_HD int mg_kind_space(int kind) {
  return (kind == MG_2D_TM || kind == MG_2D_TE) ? 2 : 3;
}
