// Device helpers shared by the forward and adjoint kernels: the tap update,
// the run correlations and the LDS tile staging.
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
// Element i of a read-only table through the constant address space, the
// index taken as an unsigned 32-bit byte offset: the scalar load adds it to
// the table's base itself, with no 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *const_at(const T *base, unsigned byte_off)
{
    return reinterpret_cast<const __attribute__((address_space(4))) T *>(
        reinterpret_cast<const __attribute__((address_space(4))) char *>(as_const(base)) + byte_off);
}
__device__ __forceinline__ float load_at(const float *base, unsigned i) { return *const_at(base, 4u * i); }
__device__ __forceinline__ int2 load_at(const int2 *base, unsigned i)
{
    const auto q = const_at(reinterpret_cast<const int *>(base), 8u * i);
    return make_int2(q[0], q[1]);
}
__device__ __forceinline__ int4 load_at(const int4 *base, unsigned i)
{
    const auto q = const_at(reinterpret_cast<const int *>(base), 16u * i);
    return make_int4(q[0], q[1], q[2], q[3]);
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
// for the P pixels starting at `lane` (the LDS address of the lane's first
// pixel), with the sliding-window body (BODY 0).  Per pixel the taps are
// added in run order, then in order along the run: the row-major order of
// the dense kernel.  Weights and run records are wave-uniform scalar loads.
template <int P, bool EXACT>
__device__ __forceinline__ void correlate_window(const float *__restrict__ lane, int lp, const int4 *__restrict__ runs,
                                                 const float *__restrict__ weights, int2 kr, float (&acc)[P])
{
    for (int r = 0; r < kr.y; ++r) {
        const int4 run = load_const(runs + kr.x + r);
        const int off = __builtin_amdgcn_readfirstlane(run.x * lp + run.y);
        run_window<P, EXACT>(lane + off, weights + run.w, run.z, acc);
    }
}

// The chain body's steps, nested so that each one is entered only when its
// predecessor did not end the run: reads I and I + 1 of the window, then the
// next pair while values remain; tap J for the P pixels, then the next tap
// while taps remain.
template <int I, int NV>
__device__ __forceinline__ void chain_reads(const float *__restrict__ src, int nv, float (&v)[NV])
{
    v[I] = src[I];
    if constexpr (I + 1 < NV) v[I + 1] = src[I + 1];
    if constexpr (I + 2 < NV)
        if (I + 2 < nv) chain_reads<I + 2, NV>(src, nv, v);
}
template <int J, int P, bool EXACT, bool WEIGHTED, int NV>
__device__ __forceinline__ void chain_taps(const float (&v)[NV], const float (&c)[kMaxRun], int len, float (&acc)[P])
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if constexpr (WEIGHTED)
            acc[p] = tap<EXACT>(acc[p], c[J], v[J + p]);
        else
            acc[p] = acc[p] + v[J + p];
    }
    if constexpr (J + 1 < kMaxRun)
        if (J + 1 < len) chain_taps<J + 1, P, EXACT, WEIGHTED, NV>(v, c, len, acc);
}

// The run correlation with no dispatch on the run length (BODY 2).  Per run, the
// L + P - 1 tile values are read two at a time into one register window,
// leaving the reads once they are all in; the kMaxRun weights are one scalar
// load (the weight table is padded by kMaxRun words); then the taps go in
// order, leaving after tap L.  Every length runs the same code on the same
// registers, so the sums are never copied between the bodies of different
// lengths, as the cases of a switch on the length would.  Per pixel the taps
// are added in run order, then in order along the run: the row-major order
// of the dense kernel.  An odd L + P - 1 reads one value past
// the run's window: the next word of the same LDS allocation, never used.
typedef float Words16 __attribute__((ext_vector_type(16), aligned(4)));
static_assert(kMaxRun == 32, "the chain body loads its weights as two 16-word vectors");

template <int P, bool EXACT, bool WEIGHTED>
__device__ __forceinline__ void correlate_chain(const float *__restrict__ lane, const int4 *__restrict__ runs_b,
                                                const float *__restrict__ weights, int2 kr, float (&acc)[P])
{
    constexpr int NV = kMaxRun + P - 1;
    const char *__restrict__ base = reinterpret_cast<const char *>(lane);
    for (int r = 0; r < kr.y; ++r) {
        // {byte offset in the tile, length, byte offset of the first weight}:
        // no address arithmetic on the scalar unit per run
        const int4 run = load_at(runs_b, (unsigned)(kr.x + r));
        __builtin_assume(run.y >= 1);  // every run has a tap: the first P reads need no exit test
        const float *__restrict__ src =
            reinterpret_cast<const float *>(base + __builtin_amdgcn_readfirstlane(run.x));
        float c[kMaxRun];
        if constexpr (WEIGHTED) {
            // two 16-word loads: a load per tap would be sunk into the tap
            // chain, each with its own wait.  The offset is an unsigned
            // 32-bit byte count, so the loads add it to the table's base
            // themselves.
            const auto cw = reinterpret_cast<const __attribute__((address_space(4))) Words16 *>(
                reinterpret_cast<const __attribute__((address_space(4))) char *>(as_const(weights)) +
                (unsigned)run.z);
            const Words16 lo = cw[0], hi = cw[1];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                c[j] = lo[j];
                c[16 + j] = hi[j];
            }
        }
        float v[NV];
        chain_reads<0, NV>(src, run.y + P - 1, v);
        chain_taps<0, P, EXACT, WEIGHTED, NV>(v, c, run.y, acc);
    }
}

// Weighted correlation of kernel k of a set: BODY 2, the chain body; BODY 0,
// the sliding window.
template <int P, bool EXACT, int BODY = 2>
__device__ __forceinline__ void correlate(const float *__restrict__ lane, int lp, const DevKset &ks, int k,
                                          float (&acc)[P])
{
    if constexpr (BODY == 2)
        correlate_chain<P, EXACT, true>(lane, ks.runs_b, ks.weights, load_at(ks.kern, (unsigned)k), acc);
    else
        correlate_window<P, EXACT>(lane, lp, ks.runs, ks.weights, load_const(ks.kern + k), acc);
}

// Unweighted sums over the runs kr (the float ring walk's ring sums), the
// runs as byte records for the tile's pitch (DevRings::runs_b).
template <int P>
__device__ __forceinline__ void ring_sums(const float *__restrict__ lane, const int4 *__restrict__ runs_b, int2 kr,
                                          float (&acc)[P])
{
    correlate_chain<P, false, false>(lane, runs_b, nullptr, kr, acc);
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

// rs[p] += fixed-point sum over the runs of a kernel's ring, for the P
// pixels whose left-boundary prefix values start at `lane` (prefix table as
// written by stage_prefix): a run (row dr, column c0, length L) adds
// Q[dr][c0 + L + p] - Q[dr][c0 + p], whatever its length.  The runs come as
// byte-offset pairs (DevRings::pairs) of the table's row pitch, so one scalar
// load brings two runs' four boundaries and no scalar arithmetic forms them;
// the four windows of P prefix values are read together (4P reads in
// flight), and the pair costs three integer operations per pixel.  There is
// no branch on the run length.  kp = {first pair, number of runs}.
template <int P>
__device__ __forceinline__ void prefix_pairs(const uint32_t *__restrict__ lane, const int4 *__restrict__ pairs,
                                             int2 kp, uint32_t (&rs)[P])
{
    const char *__restrict__ base = reinterpret_cast<const char *>(lane);
    unsigned pp = (unsigned)kp.x;  // the pair's index in the table, loaded by byte offset
    int r = 0;
    for (; r + 1 < kp.y; r += 2, ++pp) {
        const int4 o = load_at(pairs, pp);
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
        const int2 o = load_at(reinterpret_cast<const int2 *>(pairs), 2u * pp);
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
