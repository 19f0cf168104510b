// Complex FP64 matrix products on the gfx950 matrix cores
// (v_mfma_f64_16x16x4_f64, FP64 multiply and accumulate).
//
// One wavefront computes a 16 x 16 tile D += A B with K in steps of 4. Lane l
// (g = l / 16, c = l % 16) supplies A[c][k0 + g] and B[k0 + g][c]; it holds
// the result D[g + 4 i][c], i = 0..3 (the FP64 output layout: the four lane
// groups interleave the tile rows). A complex product takes four real MFMAs:
// Dr += Ar Br - Ai Bi, Di += Ar Bi + Ai Br.
#pragma once
#include <hip/hip_runtime.h>

namespace tq {

typedef double mfma_v4d __attribute__((ext_vector_type(4)));

// Output row of element i held by lane group g.
__device__ __forceinline__ constexpr int mfma_out_row(int g, int i) { return g + 4 * i; }

// Complex accumulator of one 16 x 16 tile (per lane: 4 elements).
struct CTile {
    mfma_v4d re, im;
    __device__ __forceinline__ void zero() {
        re = mfma_v4d{0.0, 0.0, 0.0, 0.0};
        im = mfma_v4d{0.0, 0.0, 0.0, 0.0};
    }
    // += a * b for this lane's operands a = A[c][k], b = B[k][c] of one K step of 4
    __device__ __forceinline__ void mac(double2 a, double2 b) {
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re, 0, 0, 0);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, re, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, im, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im, 0, 0, 0);
    }
    __device__ __forceinline__ double2 get(int i) const { return make_double2(re[i], im[i]); }
};

}  // namespace tq
