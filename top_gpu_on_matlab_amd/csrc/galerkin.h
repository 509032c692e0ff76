// Galerkin coarse-stiffness core shared by the MEX entry points and the device solver:
//   Kout = sum_s  B_s^T * Kch_s * B_s      (24 x 24, FP32/FP64)
// One 64-lane wavefront per coarse element; lane t owns the 3x3 output block
// (rows 3*(t%8).., cols 3*(t/8)..).  B (row-major), Kch (column-major) and T = Kch*B
// (row-major) are staged in LDS.  Every sum runs over k = 0..23 in order, and each child's
// contribution is formed before it is added to the accumulator.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

constexpr int GK_THREADS = 64;

// Src must provide:
//   __device__ int child(int s) const;              // < 0: skip child s
//   __device__ T kch(int f, int i) const;            // entry i (= r + 24c) of the child's matrix
template <typename T, class Src>
__device__ inline void galerkin_block(T acc[9], const Src &src, int nCh, const T *__restrict__ P,
                                      T *sB, T *sK, T *sT) {
    const int t = threadIdx.x;
    const int rb = 3 * (t & 7), cb = 3 * (t >> 3);
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = (T)0;
    for (int s = 0; s < nCh; ++s) {
        const int f = src.child(s);
        if (f < 0) continue;  // uniform over the wavefront
        __syncthreads();
        for (int i = t; i < 576; i += GK_THREADS) {
            const int r = i % 24, c = i / 24;
            sB[r * 24 + c] = P[(size_t)s * 576 + i];
            sK[i] = src.kch(f, i);
        }
        __syncthreads();
        T tb[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) tb[i] = (T)0;
        for (int k = 0; k < 24; ++k) {
            const T k0 = sK[rb + 0 + 24 * k], k1 = sK[rb + 1 + 24 * k], k2 = sK[rb + 2 + 24 * k];
            const T b0 = sB[k * 24 + cb + 0], b1 = sB[k * 24 + cb + 1], b2 = sB[k * 24 + cb + 2];
            tb[0] += k0 * b0; tb[1] += k0 * b1; tb[2] += k0 * b2;
            tb[3] += k1 * b0; tb[4] += k1 * b1; tb[5] += k1 * b2;
            tb[6] += k2 * b0; tb[7] += k2 * b1; tb[8] += k2 * b2;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) sT[(rb + i) * 24 + cb + j] = tb[3 * i + j];
        __syncthreads();
        T ub[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) ub[i] = (T)0;
        for (int k = 0; k < 24; ++k) {
            const T b0 = sB[k * 24 + rb + 0], b1 = sB[k * 24 + rb + 1], b2 = sB[k * 24 + rb + 2];
            const T t0 = sT[k * 24 + cb + 0], t1 = sT[k * 24 + cb + 1], t2 = sT[k * 24 + cb + 2];
            ub[0] += b0 * t0; ub[1] += b0 * t1; ub[2] += b0 * t2;
            ub[3] += b1 * t0; ub[4] += b1 * t1; ub[5] += b1 * t2;
            ub[6] += b2 * t0; ub[7] += b2 * t1; ub[8] += b2 * t2;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += ub[i];
    }
}

// Output entry index (row r, col c) of lane t's block element i (0..8).
__device__ inline int galerkin_entry(int i) {
    const int t = threadIdx.x;
    const int r = 3 * (t & 7) + i / 3, c = 3 * (t >> 3) + i % 3;
    return r + 24 * c;
}

}  // namespace tg
