// Device helpers shared by the forward and adjoint kernels: the tap update,
// the sliding-window run correlation and the LDS tile staging.
#ifndef LPA_TILE_HPP
#define LPA_TILE_HPP

#include <hip/hip_runtime.h>

#include "lpa_device.hpp"

namespace lpa {

// Separately rounded multiply and add: the same bits as the CPU path.
__device__ __forceinline__ float mul_add_rounded(float acc, float a, float b)
{
#pragma clang fp contract(off)
    return acc + a * b;
}

template <bool EXACT>
__device__ __forceinline__ float tap(float acc, float w, float v)
{
    if constexpr (EXACT)
        return mul_add_rounded(acc, w, v);
    else
        return __builtin_fmaf(w, v, acc);
}

// acc[p] += sum over kernel k's runs of w[j] * tile[row, col + p + j], for the
// P pixels starting at `lane` (the LDS address of the lane's first pixel).
// Per pixel the taps are added in run order, then in order along the run:
// the row-major order of the dense kernel.  A window of P+3 tile values slides
// along the run four taps at a time, so each LDS value read feeds up to P
// FMAs; weights and run records are wave-uniform scalar loads.
template <int P, bool EXACT>
__device__ __forceinline__ void correlate(const float *__restrict__ lane, int lp, const DevKset &ks, int k,
                                          float (&acc)[P])
{
    const int2 kr = ks.kern[k];
    for (int r = 0; r < kr.y; ++r) {
        const int4 run = ks.runs[kr.x + r];
        const float *__restrict__ src = lane + run.x * lp + run.y;
        const float *__restrict__ w = ks.weights + run.w;
        const int len = run.z;
        float v[P + 3];
#pragma unroll
        for (int q = 0; q < P - 1; ++q) v[q] = src[q];
        int j = 0;
        for (; j + 4 <= len; j += 4) {
#pragma unroll
            for (int q = P - 1; q < P + 3; ++q) v[q] = src[j + q];
            const float w0 = w[j], w1 = w[j + 1], w2 = w[j + 2], w3 = w[j + 3];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                acc[p] = tap<EXACT>(acc[p], w0, v[p]);
                acc[p] = tap<EXACT>(acc[p], w1, v[p + 1]);
                acc[p] = tap<EXACT>(acc[p], w2, v[p + 2]);
                acc[p] = tap<EXACT>(acc[p], w3, v[p + 3]);
            }
#pragma unroll
            for (int q = 0; q < P - 1; ++q) v[q] = v[q + 4];
        }
        for (; j < len; ++j) {
            v[P - 1] = src[j + P - 1];
            const float wj = w[j];
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] = tap<EXACT>(acc[p], wj, v[p]);
#pragma unroll
            for (int q = 0; q < P - 1; ++q) v[q] = v[q + 1];
        }
    }
}

// Stage the (TW + 2R) x (TH + 2R) window whose top-left image pixel is
// (x0 - R, y0 - R) into `tile` (row pitch lp); pixels outside the image read
// as zero.  Rows of 64 lanes read 64 consecutive floats.  No barrier.
template <int TW, int TH>
__device__ __forceinline__ void stage_tile(float *__restrict__ tile, int lp, const float *__restrict__ x, int W, int H,
                                           int pitch, int x0, int y0, int R)
{
    const int tid = threadIdx.x;
    const int fw = TW + 2 * R, fh = TH + 2 * R;
    for (int r = tid >> 6; r < fh; r += kThreads / 64) {
        const int gy = y0 - R + r;
        const bool row_in = gy >= 0 && gy < H;
        const float *__restrict__ xrow = x + (size_t)(row_in ? gy : 0) * pitch;
        for (int c = tid & 63; c < fw; c += 64) {
            const int gx = x0 - R + c;
            tile[r * lp + c] = (row_in && gx >= 0 && gx < W) ? xrow[gx] : 0.0f;
        }
    }
}

}  // namespace lpa

#endif
