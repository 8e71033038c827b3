// Device helpers shared by the forward and adjoint kernels: the tap update,
// the sliding-window run correlation and the LDS tile staging.
#ifndef LPA_TILE_HPP
#define LPA_TILE_HPP

#include <hip/hip_runtime.h>

#include "lpa_device.hpp"

namespace lpa {

// Pointers read out of descriptors are generic; these casts let the compiler
// use global (not flat) memory instructions, which do not tie up the LDS
// wait counter.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *as_global(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// Separately rounded multiply and add: the same bits as the CPU path.
__device__ __forceinline__ float mul_add_rounded(float acc, float a, float b)
{
#pragma clang fp contract(off)
    return acc + a * b;
}

// max(a, b) and min(a, b) as single VALU instructions.  fmaxf/fminf make the
// compiler quiet signalling NaNs on loop-carried operands first (one extra
// v_max per operand); the values here are arithmetic results, never
// signalling NaNs, so the plain instruction gives the same result.
__device__ __forceinline__ float max_f32(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float min_f32(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <bool EXACT>
__device__ __forceinline__ float tap(float acc, float w, float v)
{
    if constexpr (EXACT)
        return mul_add_rounded(acc, w, v);
    else
        return __builtin_fmaf(w, v, acc);
}

// One run of exactly L taps for the P pixels starting at `src`: L + P - 1
// tile values and L weights, then L*P FMAs.  Per pixel the taps are added in
// order along the run.
template <int L, int P, bool EXACT, bool WEIGHTED = true>
__device__ __forceinline__ void run_taps(const float *__restrict__ src, const float *__restrict__ w, float (&acc)[P])
{
    float v[L + P - 1];
#pragma unroll
    for (int q = 0; q < L + P - 1; ++q) v[q] = src[q];
    if constexpr (WEIGHTED) {
        float c[L];
#pragma unroll
        for (int j = 0; j < L; ++j) c[j] = w[j];
#pragma unroll
        for (int j = 0; j < L; ++j)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] = tap<EXACT>(acc[p], c[j], v[j + p]);
    } else {
        (void)w;
#pragma unroll
        for (int j = 0; j < L; ++j)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] = acc[p] + v[j + p];
    }
}

// Sliding-window body: a window of P+3 tile values moves along the run four
// taps at a time, then one tap at a time for the rest.  Smaller code than the
// unrolled body; costs window moves and a loop tail.
template <int P, bool EXACT>
__device__ __forceinline__ void run_window(const float *__restrict__ src, const float *__restrict__ w, int len,
                                           float (&acc)[P])
{
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

// Runs are at most kMaxRun taps long (the packer splits longer ones).
constexpr int kMaxRun = 32;

// acc[p] += sum over the runs kr = {first, count} of w[j] * tile[row, col + p + j]
// (or of tile[row, col + p + j] alone when !WEIGHTED), for the P pixels
// starting at `lane` (the LDS address of the lane's first pixel).  Per pixel
// the taps are added in run order, then in order along the run: the
// row-major order of the dense kernel.  Each LDS value read feeds up to P
// FMAs; the run's address costs one add from the lane's base; weights and
// run records are wave-uniform scalar loads.  BODY 1: the run length is the
// same for every lane, so a switch picks one fully unrolled body per run (no
// window shuffling, no loop tail).  BODY 0: the sliding-window loop.
template <int P, bool EXACT, int BODY, bool WEIGHTED>
__device__ __forceinline__ void correlate_runs(const float *__restrict__ lane, int lp, const int4 *__restrict__ runs,
                                               const float *__restrict__ weights, int2 kr, float (&acc)[P])
{
    for (int r = 0; r < kr.y; ++r) {
        const int4 run = runs[kr.x + r];
        const int off = __builtin_amdgcn_readfirstlane(run.x * lp + run.y);
        const float *__restrict__ src = lane + off;
        const float *__restrict__ w = weights + run.w;
        if constexpr (BODY == 0 && WEIGHTED) {
            run_window<P, EXACT>(src, w, run.z, acc);
            continue;
        }
        switch (run.z) {
#define LPA_RUN_CASE(L)                                                                                            \
    case L: run_taps<L, P, EXACT, WEIGHTED>(src, w, acc); break;
            LPA_RUN_CASE(1) LPA_RUN_CASE(2) LPA_RUN_CASE(3) LPA_RUN_CASE(4) LPA_RUN_CASE(5) LPA_RUN_CASE(6)
            LPA_RUN_CASE(7) LPA_RUN_CASE(8) LPA_RUN_CASE(9) LPA_RUN_CASE(10) LPA_RUN_CASE(11) LPA_RUN_CASE(12)
            LPA_RUN_CASE(13) LPA_RUN_CASE(14) LPA_RUN_CASE(15) LPA_RUN_CASE(16) LPA_RUN_CASE(17) LPA_RUN_CASE(18)
            LPA_RUN_CASE(19) LPA_RUN_CASE(20) LPA_RUN_CASE(21) LPA_RUN_CASE(22) LPA_RUN_CASE(23) LPA_RUN_CASE(24)
            LPA_RUN_CASE(25) LPA_RUN_CASE(26) LPA_RUN_CASE(27) LPA_RUN_CASE(28) LPA_RUN_CASE(29) LPA_RUN_CASE(30)
            LPA_RUN_CASE(31) LPA_RUN_CASE(32)
#undef LPA_RUN_CASE
        default: break;
        }
    }
}

// Weighted correlation of kernel k of a set (see correlate_runs).
template <int P, bool EXACT, int BODY = 1>
__device__ __forceinline__ void correlate(const float *__restrict__ lane, int lp, const DevKset &ks, int k,
                                          float (&acc)[P])
{
    correlate_runs<P, EXACT, BODY, true>(lane, lp, ks.runs, ks.weights, ks.kern[k], acc);
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
    const auto gx = as_global(x);
    for (int r = tid >> 6; r < fh; r += kThreads / 64) {
        const int gy = y0 - R + r;
        const bool row_in = gy >= 0 && gy < H;
        const auto xrow = gx + (size_t)(row_in ? gy : 0) * pitch;
        for (int c = tid & 63; c < fw; c += 64) {
            const int gx = x0 - R + c;
            tile[r * lp + c] = (row_in && gx >= 0 && gx < W) ? xrow[gx] : 0.0f;
        }
    }
}

}  // namespace lpa

#endif
