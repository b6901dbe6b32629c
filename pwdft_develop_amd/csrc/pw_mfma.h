// FP64 MFMA helpers shared by the dense kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace pw {

typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma_f64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Lane src's value of v, in every lane (src uniform).
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Accumulator layout of v_mfma_f64_16x16x4f64 for a 32 x 32 block held as
// four 16 x 16 tiles q = ti + 2 tj: lane l, register i holds element
// (16 ti + 4 i + l / 16, 16 tj + l % 16).  Register i of a tile is also the
// B operand of the tile's k-step 4 i (B lane l = (k l / 16, n l % 16)).
// Row-major LDS copies of a block:
__device__ __forceinline__ void blk_to_lds(const d4v (&a)[4], double* M, int ldm, int lane) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) M[(16 * (q & 1) + 4 * i + (lane >> 4)) * ldm + 16 * (q >> 1) + (lane & 15)] = a[q][i];
}
__device__ __forceinline__ void blk_from_lds(d4v (&a)[4], const double* M, int ldm, int lane) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) a[q][i] = M[(16 * (q & 1) + 4 * i + (lane >> 4)) * ldm + 16 * (q >> 1) + (lane & 15)];
}

}  // namespace pw
