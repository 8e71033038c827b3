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

// Read-only tables (kernel runs and weights, per-image tables, descriptors,
// tile lists) never change while a kernel runs.  Reading them through the
// constant address space lets the compiler keep wave-uniform reads on the
// scalar unit even after the kernel's own global stores and atomics.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *as_const(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}
// The same for HIP's vector types, whose constructors take generic references.
__device__ __forceinline__ int4 load_const(const int4 *p)
{
    const auto q = as_const((const int *)p);
    return make_int4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ int2 load_const(const int2 *p)
{
    const auto q = as_const((const int *)p);
    return make_int2(q[0], q[1]);
}
__device__ __forceinline__ uint2 load_const(const uint2 *p)
{
    const auto q = as_const((const unsigned *)p);
    return make_uint2(q[0], q[1]);
}
// A descriptor struct, word by word.
template <typename T>
__device__ __forceinline__ T load_const_struct(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "descriptor size must be whole words");
    const auto q = as_const((const unsigned *)p);
    unsigned w[sizeof(T) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
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
        for (int j = 0; j < L; ++j) c[j] = as_const(w)[j];
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
        const auto cw = as_const(w);
        const float w0 = cw[j], w1 = cw[j + 1], w2 = cw[j + 2], w3 = cw[j + 3];
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
        const float wj = as_const(w)[j];
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
        const int4 run = load_const(runs + kr.x + r);
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

// correlate_runs over a kernel's stretches of consecutive runs of one length
// ({first run, number of runs, length}): one dispatch on the length per
// stretch, then a loop over its runs with the unrolled body.  The taps are
// added in the same order as run by run.  Each dispatch leaves the sums in
// the registers the switch's cases share and copies them back to the loop's,
// so the fewer dispatches, the fewer copies and scalar branches.
template <int P, bool EXACT>
__device__ __forceinline__ void correlate_groups(const float *__restrict__ lane, int lp,
                                                 const int4 *__restrict__ runs, const float *__restrict__ weights,
                                                 const int4 *__restrict__ groups, int2 kg, float (&acc)[P])
{
    for (int g = 0; g < kg.y; ++g) {
        const int4 grp = load_const(groups + kg.x + g);
        switch (grp.z) {
#define LPA_GROUP_CASE(L)                                                                                          \
    case L:                                                                                                        \
        for (int r = 0; r < grp.y; ++r) {                                                                          \
            const int4 run = load_const(runs + grp.x + r);                                                         \
            const float *__restrict__ src = lane + __builtin_amdgcn_readfirstlane(run.x * lp + run.y);             \
            run_taps<L, P, EXACT, true>(src, weights + run.w, acc);                                                \
        }                                                                                                          \
        break;
            LPA_GROUP_CASE(1) LPA_GROUP_CASE(2) LPA_GROUP_CASE(3) LPA_GROUP_CASE(4) LPA_GROUP_CASE(5)
            LPA_GROUP_CASE(6) LPA_GROUP_CASE(7) LPA_GROUP_CASE(8) LPA_GROUP_CASE(9) LPA_GROUP_CASE(10)
            LPA_GROUP_CASE(11) LPA_GROUP_CASE(12) LPA_GROUP_CASE(13) LPA_GROUP_CASE(14) LPA_GROUP_CASE(15)
            LPA_GROUP_CASE(16) LPA_GROUP_CASE(17) LPA_GROUP_CASE(18) LPA_GROUP_CASE(19) LPA_GROUP_CASE(20)
            LPA_GROUP_CASE(21) LPA_GROUP_CASE(22) LPA_GROUP_CASE(23) LPA_GROUP_CASE(24) LPA_GROUP_CASE(25)
            LPA_GROUP_CASE(26) LPA_GROUP_CASE(27) LPA_GROUP_CASE(28) LPA_GROUP_CASE(29) LPA_GROUP_CASE(30)
            LPA_GROUP_CASE(31) LPA_GROUP_CASE(32)
#undef LPA_GROUP_CASE
        default: break;
        }
    }
}

// Weighted correlation of kernel k of a set (see correlate_runs): the
// unrolled body goes by stretches of runs of one length.
template <int P, bool EXACT, int BODY = 1>
__device__ __forceinline__ void correlate(const float *__restrict__ lane, int lp, const DevKset &ks, int k,
                                          float (&acc)[P])
{
    if constexpr (BODY == 1)
        correlate_groups<P, EXACT>(lane, lp, ks.runs, ks.weights, ks.groups, load_const(ks.kern_groups + k), acc);
    else
        correlate_runs<P, EXACT, BODY, true>(lane, lp, ks.runs, ks.weights, load_const(ks.kern + k), acc);
}

// Stage the (TW + 2R) x (TH + 2R) window whose top-left image pixel is
// (x0 - R, y0 - R) into `tile` (row pitch lp); pixels outside the image read
// as zero.  Rows of 64 lanes read 64 consecutive floats.  No barrier.
template <int TW, int TH, int NT = kThreads>
__device__ __forceinline__ void stage_tile(float *__restrict__ tile, int lp, const float *__restrict__ x, int W, int H,
                                           int pitch, int x0, int y0, int R)
{
    const int tid = threadIdx.x;
    const int fw = TW + 2 * R, fh = TH + 2 * R;
    const auto gx = as_global(x);
    for (int r = tid >> 6; r < fh; r += NT / 64) {
        const int gy = y0 - R + r;
        const bool row_in = gy >= 0 && gy < H;
        const auto xrow = gx + (size_t)(row_in ? gy : 0) * pitch;
        for (int c = tid & 63; c < fw; c += 64) {
            const int gx = x0 - R + c;
            tile[r * lp + c] = (row_in && gx >= 0 && gx < W) ? xrow[gx] : 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------
// Fixed-point ring sums.  For a kernel set in ring form (uniform weight w_s
// per kernel, nested supports) the estimate of scale s is w_s times the sum
// of the tile over the support.  Tile values are taken in fixed point,
// q(v) = rint(v * 2^F), so sums are exact 32-bit integers, independent of
// order, and a row run's sum is a difference of two row-prefix values:
//   sum_{c=a}^{b} q(x[r][c]) = Q[r][b + 1] - Q[r][a],  Q[r][j] = sum_{c<j} q(x[r][c]).
// F is chosen per kernel set so that any support sum of values below the
// set's range bound fits in 31 bits; a window holding a value outside the
// range (or not finite) is walked with float ring sums instead.  All integer
// arithmetic is modulo 2^32; the final sums are exact.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t quantize(float v, float qscale)
{
    return (uint32_t)(int)__builtin_rintf(v * qscale);
}

// a * b rounded once (never fused into a following add).
__device__ __forceinline__ float mul_rounded(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}

// Inclusive prefix sum over the 64 lanes of a wave: row_shr 1, 2, 4, 8
// within each row of 16 lanes, then row_bcast15 / row_bcast31 carry the row
// totals forward.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, true);
}
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    v += dpp_or_zero<0x111>(v);
    v += dpp_or_zero<0x112>(v);
    v += dpp_or_zero<0x114>(v);
    v += dpp_or_zero<0x118>(v);
    v += dpp_or_zero<0x142, 0xa>(v);
    v += dpp_or_zero<0x143, 0xc>(v);
    return v;
}

// Work-group AND of v, returned to every lane as a wave-uniform value; also
// a barrier.  Built from wave ballots and a __syncthreads: a vote through
// __syncthreads_and would act as a memory clobber for the compiler, and every
// kernel-table load after it would become a vector load.  The vote slots are
// rewritten by the next call only after the caller's next barrier.
template <int NT, bool ALL>
__device__ __forceinline__ bool block_vote(bool v)
{
    __shared__ int votes[NT / 64];
    const bool w = ALL ? __all(v) : __any(v);
    if ((threadIdx.x & 63) == 0) votes[threadIdx.x >> 6] = w;
    __syncthreads();
    int r = ALL ? 1 : 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r = ALL ? (r & votes[i]) : (r | votes[i]);
    return __builtin_amdgcn_readfirstlane(r) != 0;
}
__device__ __forceinline__ bool block_all(bool v) { return block_vote<kThreads, true>(v); }

// Work-group maximum of v, wave-uniform in every lane.  Two barriers: one
// before the slots are read, one after, so back-to-back calls do not race.
template <int NT>
__device__ __forceinline__ int block_max(int v)
{
    __shared__ int wave_max[NT / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = wave_max[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) r = max(r, wave_max[i]);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(r);
}

// Stage the (TW + 2R) x (TH + 2R) window (as stage_tile) as row-prefix sums
// of fixed-point values: q[r * lp + j] = Q[r][j] for j = 0 .. TW + 2R
// (Q[r][0] = 0).  Each wave takes whole rows, two values per lane, and scans
// them with wave_inclusive_sum.  Returns, to every thread of the work-group,
// whether every value of the window lies strictly within +-qrange (this is
// also the barrier after staging).
template <int TW, int TH, int NT = kThreads>
__device__ __forceinline__ bool stage_prefix(uint32_t *__restrict__ q, int lp, const float *__restrict__ x, int W,
                                             int H, int pitch, int x0, int y0, int R, float qscale, float qrange)
{
    static_assert(TW + 2 * 31 <= 128, "a window row must fit two values per lane");
    const int tid = threadIdx.x, lane = tid & 63;
    const int fw = TW + 2 * R, fh = TH + 2 * R;
    const auto gx = as_global(x);
    const int c = 2 * lane, gc = x0 - R + c;
    const bool in0 = c < fw && gc >= 0 && gc < W, in1 = c + 1 < fw && gc + 1 >= 0 && gc + 1 < W;
    bool fits = true;
    for (int r = tid >> 6; r < fh; r += NT / 64) {
        const int gy = y0 - R + r;
        const bool row_in = gy >= 0 && gy < H;
        const auto xrow = gx + (size_t)(row_in ? gy : 0) * pitch;
        const float v0 = row_in && in0 ? xrow[gc] : 0.0f;
        const float v1 = row_in && in1 ? xrow[gc + 1] : 0.0f;
        fits = fits && __builtin_fabsf(v0) < qrange && __builtin_fabsf(v1) < qrange;
        const uint32_t q1 = quantize(v1, qscale);
        const uint32_t incl = wave_inclusive_sum(quantize(v0, qscale) + q1);
        uint32_t *__restrict__ row = q + r * lp;
        if (c < fw) row[1 + c] = incl - q1;
        if (c + 1 < fw) row[2 + c] = incl;
        if (lane == 0) row[0] = 0u;
    }
    return block_vote<NT, true>(fits);
}

// a + b + c in one v_add3_u32 (wrapping, as the fixed-point sums are).
__device__ __forceinline__ uint32_t add3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// rs[p] += fixed-point sum over the runs kr of the tile, for the P pixels
// whose left-boundary prefix values start at `lane` (prefix table as written
// by stage_prefix): a run (row dr, column c0, length L) adds
// Q[dr][c0 + L + p] - Q[dr][c0 + p], whatever its length.  Runs go in pairs:
// one scalar load brings both records, the four windows of P prefix values
// are read together (4P reads in flight), and the pair costs three integer
// operations per pixel.  There is no branch on the run length.
template <int P>
__device__ __forceinline__ void prefix_runs(const uint32_t *__restrict__ lane, int lp, const int4 *__restrict__ runs,
                                            int2 kr, uint32_t (&rs)[P])
{
    int r = 0;
    for (; r + 1 < kr.y; r += 2) {
        const int4 u = load_const(runs + kr.x + r), v = load_const(runs + kr.x + r + 1);
        const int ou = u.x * lp + u.y, ov = v.x * lp + v.y;
        const uint32_t *__restrict__ ua = lane + ou;
        const uint32_t *__restrict__ ub = lane + (ou + u.z);
        const uint32_t *__restrict__ va = lane + ov;
        const uint32_t *__restrict__ vb = lane + (ov + v.z);
        uint32_t a0[P], b0[P], a1[P], b1[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            a0[p] = ua[p];
            b0[p] = ub[p];
            a1[p] = va[p];
            b1[p] = vb[p];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) rs[p] = add3(rs[p], b0[p], b1[p]) - a0[p] - a1[p];
    }
    if (r < kr.y) {
        const int4 u = load_const(runs + kr.x + r);
        const int ou = u.x * lp + u.y;
        const uint32_t *__restrict__ ua = lane + ou;
        const uint32_t *__restrict__ ub = lane + (ou + u.z);
        uint32_t a0[P], b0[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            a0[p] = ua[p];
            b0[p] = ub[p];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) rs[p] += b0[p] - a0[p];
    }
}

// prefix_runs with the runs as byte-offset pairs (DevRings::pairs) of the
// table's row pitch: one scalar load brings both runs' four boundaries, and
// no scalar arithmetic forms them.  kp = {first pair, number of runs}.
template <int P>
__device__ __forceinline__ void prefix_pairs(const uint32_t *__restrict__ lane, const int4 *__restrict__ pairs,
                                             int2 kp, uint32_t (&rs)[P])
{
    const char *__restrict__ base = reinterpret_cast<const char *>(lane);
    const int4 *__restrict__ pp = pairs + kp.x;
    int r = 0;
    for (; r + 1 < kp.y; r += 2, ++pp) {
        const int4 o = load_const(pp);
        const uint32_t *__restrict__ ua = reinterpret_cast<const uint32_t *>(base + o.x);
        const uint32_t *__restrict__ ub = reinterpret_cast<const uint32_t *>(base + o.y);
        const uint32_t *__restrict__ va = reinterpret_cast<const uint32_t *>(base + o.z);
        const uint32_t *__restrict__ vb = reinterpret_cast<const uint32_t *>(base + o.w);
        uint32_t a0[P], b0[P], a1[P], b1[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            a0[p] = ua[p];
            b0[p] = ub[p];
            a1[p] = va[p];
            b1[p] = vb[p];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) rs[p] = add3(rs[p], b0[p], b1[p]) - a0[p] - a1[p];
    }
    if (r < kp.y) {
        const int2 o = load_const(reinterpret_cast<const int2 *>(pp));
        const uint32_t *__restrict__ ua = reinterpret_cast<const uint32_t *>(base + o.x);
        const uint32_t *__restrict__ ub = reinterpret_cast<const uint32_t *>(base + o.y);
        uint32_t a0[P], b0[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            a0[p] = ua[p];
            b0[p] = ub[p];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) rs[p] += b0[p] - a0[p];
    }
}

// The same fixed-point sums from a float tile (quantizing each value read):
// rq[p] += sum over the runs kr of q(tile[row, col + p + j]).
template <int L, int P>
__device__ __forceinline__ void run_taps_q(const float *__restrict__ src, float qscale, uint32_t (&rq)[P])
{
    uint32_t v[L + P - 1];
#pragma unroll
    for (int i = 0; i < L + P - 1; ++i) v[i] = quantize(src[i], qscale);
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
        for (int p = 0; p < P; ++p) rq[p] += v[j + p];
}
template <int P>
__device__ __forceinline__ void quantized_runs(const float *__restrict__ lane, int lp, const int4 *__restrict__ runs,
                                               int2 kr, float qscale, uint32_t (&rq)[P])
{
    for (int r = 0; r < kr.y; ++r) {
        const int4 run = load_const(runs + kr.x + r);
        const int off = __builtin_amdgcn_readfirstlane(run.x * lp + run.y);
        const float *__restrict__ src = lane + off;
        switch (run.z) {
#define LPA_QRUN_CASE(L)                                                                                           \
    case L: run_taps_q<L, P>(src, qscale, rq); break;
            LPA_QRUN_CASE(1) LPA_QRUN_CASE(2) LPA_QRUN_CASE(3) LPA_QRUN_CASE(4) LPA_QRUN_CASE(5) LPA_QRUN_CASE(6)
            LPA_QRUN_CASE(7) LPA_QRUN_CASE(8) LPA_QRUN_CASE(9) LPA_QRUN_CASE(10) LPA_QRUN_CASE(11)
            LPA_QRUN_CASE(12) LPA_QRUN_CASE(13) LPA_QRUN_CASE(14) LPA_QRUN_CASE(15) LPA_QRUN_CASE(16)
            LPA_QRUN_CASE(17) LPA_QRUN_CASE(18) LPA_QRUN_CASE(19) LPA_QRUN_CASE(20) LPA_QRUN_CASE(21)
            LPA_QRUN_CASE(22) LPA_QRUN_CASE(23) LPA_QRUN_CASE(24) LPA_QRUN_CASE(25) LPA_QRUN_CASE(26)
            LPA_QRUN_CASE(27) LPA_QRUN_CASE(28) LPA_QRUN_CASE(29) LPA_QRUN_CASE(30) LPA_QRUN_CASE(31)
            LPA_QRUN_CASE(32)
#undef LPA_QRUN_CASE
        default: break;
        }
    }
}

// stage_tile, also returning to every thread whether every staged value lies
// strictly within +-qrange (this is also the barrier after staging).
template <int TW, int TH, int NT = kThreads>
__device__ __forceinline__ bool stage_tile_fits(float *__restrict__ tile, int lp, const float *__restrict__ x, int W,
                                                int H, int pitch, int x0, int y0, int R, float qrange)
{
    const int tid = threadIdx.x;
    const int fw = TW + 2 * R, fh = TH + 2 * R;
    const auto gx = as_global(x);
    bool fits = true;
    for (int r = tid >> 6; r < fh; r += NT / 64) {
        const int gy = y0 - R + r;
        const bool row_in = gy >= 0 && gy < H;
        const auto xrow = gx + (size_t)(row_in ? gy : 0) * pitch;
        for (int c = tid & 63; c < fw; c += 64) {
            const int gc = x0 - R + c;
            const float v = (row_in && gc >= 0 && gc < W) ? xrow[gc] : 0.0f;
            fits = fits && __builtin_fabsf(v) < qrange;
            tile[r * lp + c] = v;
        }
    }
    return block_vote<NT, true>(fits);
}

}  // namespace lpa

#endif
