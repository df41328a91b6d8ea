// FP64 matrix-core helpers for gfx950 (v_mfma_f64_16x16x4f64).
//
// Lane layout of one 16x16x4 step (l = lane 0..63, i = l % 16, q = l / 16):
//   A (16 x 4): lane holds A[i][q]
//   B (4 x 16): lane holds B[q][i]
//   D (16 x 16): lane holds D[4 r + q][i] in slot r = 0..3
// (checked on the device with an in-bounds probe of known operands).
// Complex products use four real steps on split real/imaginary operands.
#pragma once
#include <hip/hip_runtime.h>

typedef double msp_d4 __attribute__((ext_vector_type(4)));

__device__ inline msp_d4 mfma_f64(double a, double b, msp_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// (Dr, Di) += (ar + i ai)(br + i bi) over one k-step
__device__ inline void mfma_cmac(double ar, double ai, double br, double bi, msp_d4 &Dr, msp_d4 &Di)
{
    Dr = mfma_f64(ar, br, Dr);
    Dr = mfma_f64(-ai, bi, Dr);
    Di = mfma_f64(ar, bi, Di);
    Di = mfma_f64(ai, br, Di);
}

__device__ inline msp_d4 d4zero()
{
    msp_d4 z = {0.0, 0.0, 0.0, 0.0};
    return z;
}

// XCD-aware tile order for grids of `ntile` tiles per instance.  Workgroups
// go to the eight XCDs round-robin by linear id L (x fastest), so XCD L % 8
// receives L, L + 8, L + 16, ...; those are handed the tiles of one instance
// in turn, and the tiles of an instance run on one XCD at about the same
// time and share its L2 (operands read by every tile: U, the V T blocks).
// The last batch % 8 instances keep the plain order.  Any order is correct;
// only the cache reuse depends on it.
__device__ inline void msp_xcd_tile(int ntile, int batch, int &tile, int &inst)
{
    const int L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int full = (batch / 8) * 8 * ntile;
    if (L < full) {
        const int slot = L >> 3;
        inst = (slot / ntile) * 8 + (L & 7);
        tile = slot - (slot / ntile) * ntile;
    } else {
        inst = L / ntile;
        tile = L - inst * ntile;
    }
}
