// Fused LPA-ICI forward for gfx950: one work-group owns an output tile,
// stages the tile plus its halo in LDS once, and every lane walks the scales
// of all directions for `px` adjacent pixels with the ICI intervals, the
// selected estimate and the inverse-variance aggregation kept in registers.
// Only the denoised image (and, optionally, a uint8 scale index per pixel and
// direction) is written to HBM.
//
// Behaviour restated from the reference:
//   tap order and zero padding   src/v0_0_lpa_ici_2D_naive.cpp:785-814
//   ICI update, non-strict U>=L   src/v0_0_lpa_ici_2D_naive.cpp:889-903
//   early exit on an empty intersection  src/v4_1_lpa_ici_2D_sparse_shared.cu:1077-1087
//   aggregation order over d      src/v0_0_lpa_ici_2D_naive.cpp:917-927
// The reference runs 16 walk launches + 16 accumulate launches + 1 normalize
// launch per image and refills its LDS tile once per direction
// (src/v4_1_lpa_ici_2D_sparse_shared.cu:1219-1235); here one launch covers
// every direction and the tile is filled once.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lpa_device.hpp"
#include "lpa_tile.hpp"

namespace lpa {
namespace {

typedef float float4_words __attribute__((ext_vector_type(4)));

// The P values of a lane at consecutive addresses: as P / 4 words of four
// elements when `whole` (every pixel inside, the row start aligned to four
// elements), else element by element for the pixels inside the image.
template <int P, typename T, typename G>
__device__ __forceinline__ void store_row(G *dst, const T (&v)[P], const bool (&inside)[P], bool whole)
{
    if constexpr (P % 4 == 0) {
        if (whole) {
            using Word = typename std::conditional<sizeof(T) == 1, uint32_t, float4_words>::type;
            using GWord = __attribute__((address_space(1))) Word;
#pragma unroll
            for (int q = 0; q < P / 4; ++q) {
                Word w;
                if constexpr (sizeof(T) == 1)
                    w = (uint32_t)v[4 * q] | (uint32_t)v[4 * q + 1] << 8 | (uint32_t)v[4 * q + 2] << 16 |
                        (uint32_t)v[4 * q + 3] << 24;
                else
                    w = float4_words{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
                reinterpret_cast<GWord *>(dst)[q] = w;
            }
            return;
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
        if (inside[p]) dst[p] = v[p];
}

// The scale walk of every direction for the P pixels of one lane, then the
// aggregation and the stores.  `src` is the lane's LDS address: the float
// tile value of its first pixel (kWalkExact, kWalkFma, kWalkRingFloat) or
// the prefix value at its first pixel's column (kWalkRing, fixed point).
template <int P, int MODE, int BODY, typename T>
__device__ __forceinline__ void walk_lane(const T *__restrict__ src, int lp, const DevKset &ks, const DevRings &rings,
                                          const float *__restrict__ tab, const float *__restrict__ inv_lds,
                                          float *__restrict__ y, uint8_t *__restrict__ map,
                                          float *__restrict__ est, int W, int H, int pitch, int px0, int py)
{
    constexpr bool EXACT = MODE == kWalkExact;
    const int S = ks.S, K = ks.D * ks.S;
    const float *__restrict__ thr_t = tab + K;

    bool inside[P];
    float num[P], den[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        inside[p] = py < H && px0 + p < W;
        num[p] = 0.0f;
        den[p] = 0.0f;
    }
    const size_t plane = (size_t)W * H;
    const size_t pix = (size_t)py * W + px0;
    // Every pixel of the lane inside, and rows of the maps (and estimates)
    // and of y starting on 4-element boundaries: the stores go out as whole
    // 4-element words (px0 is a multiple of P, itself a multiple of 4).
    const bool whole4 = P % 4 == 0 && inside[P - 1];
    const bool map_aligned = W % 4 == 0 && ((uintptr_t)map & 3) == 0;
    const bool est_aligned = W % 4 == 0 && ((uintptr_t)est & 15) == 0;
    const bool y_aligned = pitch % 4 == 0 && ((uintptr_t)y & 15) == 0;

    // A wave wholly outside the image has nothing to walk or store.
    // inside_mask[p]: the wave's lanes whose pixel p is inside (wave-uniform)
    uint64_t inside_mask[P], any_inside = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        inside_mask[p] = __builtin_amdgcn_ballot_w64(inside[p]);
        any_inside |= inside_mask[p];
    }
    if (any_inside == 0) return;

    for (int d = 0; d < ks.D; ++d) {
        float lo[P], hi[P], best[P];
        int sel[P];           // the last scale whose interval is non-empty
        float ring_sum[P];    // float ring sums: running sum over scales 0..s
        uint32_t ring_q[P];   // fixed-point ring sums
#pragma unroll
        for (int p = 0; p < P; ++p) {
            ring_sum[p] = 0.0f;
            ring_q[p] = 0u;
        }
        // Scale 0: the interval [yhat - thr, yhat + thr] with thr >= 0 is
        // never empty (a NaN bound leaves the infinite start in place, as
        // the reference's std::max / std::min do), so every pixel selects
        // it.  The fixed-point walk takes it without the test; the other
        // walks (whose estimate code is long) run it through the loop.
        constexpr int PEEL = MODE == kWalkRing ? 1 : 0;
        if constexpr (PEEL) {
            // yhat_0 = (w_0 2^-F) * (fixed-point sum over the support of
            // scale 0)
            prefix_pairs<P>(src, rings.pairs, load_at(rings.pair_kern, (unsigned)(d * S)), ring_q);
            const float wq = load_at(rings.scale_q, (unsigned)(d * S));
            const float thr = load_at(thr_t, (unsigned)(d * S));
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float acc = mul_rounded((float)(int)ring_q[p], wq);
                lo[p] = max_f32(-INFINITY, acc - thr);
                hi[p] = min_f32(INFINITY, acc + thr);
                best[p] = acc;
                sel[p] = 0;
            }
        } else {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                lo[p] = -INFINITY;
                hi[p] = INFINITY;
                best[p] = 0.0f;
                sel[p] = 0;
            }
        }
        // live: the wave's lanes with a pixel inside the image whose interval
        // is still non-empty (the `ok` of the previous scale; an empty
        // intersection stays empty).  The wave stops walking this direction
        // once no lane is live.  Every lane of a walking wave walks, as a
        // lane masked off would cost the same issue cycles: a pixel whose
        // interval is empty keeps it empty and its selection, and pixels
        // outside the image keep no lane live and store nothing.
        uint64_t live = any_inside;
        for (int s = PEEL; s < S; ++s) {
            if (live == 0) break;
            const int k = d * S + s;
            // the scale's table entries are read before the lane-divergent
            // walk: read after it, their index is no longer known to be
            // wave-uniform and the reads leave the scalar unit
            const float thr = load_at(thr_t, (unsigned)k);
            float acc[P];
            if constexpr (MODE == kWalkRing) {
                // yhat_s = (w_s 2^-F) * (fixed-point sum over the support of
                // scale s), the previous scale's sum plus this scale's ring;
                // one rounding for the product, none hidden in the interval.
                const float wq = load_at(rings.scale_q, (unsigned)k);
                prefix_pairs<P>(src, rings.pairs, load_at(rings.pair_kern, (unsigned)k), ring_q);
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = mul_rounded((float)(int)ring_q[p], wq);
            } else if constexpr (MODE == kWalkRingFloat) {
                // yhat_s = w_s * (float sum over the support of scale s).
                const float w = as_const(rings.scale)[k];
                ring_sums<P>(src, rings.runs_b, load_at(rings.kern, (unsigned)k), ring_sum);
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = ring_sum[p] * w;
            } else {
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = 0.0f;
                correlate<P, EXACT, BODY>(src, lp, ks, k, acc);
            }
            uint64_t next_live = 0;  // the compares' lane masks, combined on the scalar unit
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float l = max_f32(lo[p], acc[p] - thr);
                const float h = min_f32(hi[p], acc[p] + thr);
                lo[p] = l;
                hi[p] = h;
                const bool ok = h >= l;
                if (ok) {
                    best[p] = acc[p];
                    sel[p] = s;
                }
                next_live |= __builtin_amdgcn_ballot_w64(ok) & inside_mask[p];
            }
            live = next_live;
        }
        // 1/var of the selected scale, read once per direction (the
        // intersection only shrinks, so the last scale with a non-empty
        // interval is the selected one and the walk kept its index)
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const float binv = inv_lds[d * S + sel[p]];
            num[p] = mul_add_rounded(num[p], best[p], binv);
            den[p] = den[p] + binv;
        }
        if (map) {
            uint8_t v[P];
#pragma unroll
            for (int p = 0; p < P; ++p) v[p] = (uint8_t)sel[p];
            store_row<P>(as_global(map) + (size_t)d * plane + pix, v, inside, whole4 && map_aligned);
        }
        if (est) store_row<P>(as_global(est) + (size_t)d * plane + pix, best, inside, whole4 && est_aligned);
    }
    float out[P];
#pragma unroll
    for (int p = 0; p < P; ++p) out[p] = num[p] / den[p];
    store_row<P>(as_global(y) + (size_t)py * pitch + px0, out, inside, whole4 && y_aligned);
}

// The lane's cell of a tile: column tx (in units of P pixels) and row ty.
// SPLIT: a wave covers 64 / LX rows; lanes 0-31 take the left LX / 2 lane
// columns of them, lanes 32-63 the right LX / 2.
template <int TW, int TH, int P, int NT, int SPLIT>
__device__ __forceinline__ void lane_cell(int tid, int &tx, int &ty)
{
    constexpr int LX = TW / P;
    static_assert(LX * TH == NT, "tile shape must use every lane of the work-group");
    static_assert(LX >= 2 && (LX & (LX - 1)) == 0 && LX <= 64, "lanes per row: a power of two up to 64");
    constexpr int HW = LX / 2, RW = 64 / LX;
    tx = SPLIT ? ((tid & 32) ? HW : 0) + (tid & (HW - 1)) : tid % LX;
    ty = SPLIT ? (tid >> 6) * RW + (tid & 31) / HW : tid / LX;
}

// One output tile at (x0, y0) of an image.  `tile` is the work-group's LDS
// buffer of lds_floats(TW + 1, TH, R, PM) words: the window (tile plus a halo
// of R) with one leading column, which the fixed-point form's prefix table
// needs (Q[r][0] = 0) and the float forms leave unused.  kWalkRing stages
// the window as fixed-point prefix sums and walks it if every value lies
// within the kernel set's range; otherwise, and for the other modes, the
// window is staged as floats.  Every barrier is reached by every lane.
template <int TW, int TH, int P, int MODE, int BODY, int PM, int NT, int SPLIT>
__device__ __forceinline__ void process_tile(float *__restrict__ tile, const DevKset &ks, const DevRings &rings,
                                             const float *__restrict__ tab, const float *__restrict__ x,
                                             float *__restrict__ y, uint8_t *__restrict__ map,
                                             float *__restrict__ est, int W, int H, int pitch, int x0, int y0, int R)
{
    const int lp = prefix_pitch(TW, R, PM);  // the ring pairs' offsets are in this pitch
    const int tid = threadIdx.x;
    int tx, ty;
    lane_cell<TW, TH, P, NT, SPLIT>(tid, tx, ty);
    const int px0 = x0 + tx * P, py = y0 + ty;
    // LDS offset of window column R + tx*P of row ty + R, not counting the
    // leading column.  Past the leading column (the float tile) it is the
    // lane's first pixel; in the prefix table it is Q[.][R + tx*P], the left
    // boundary of a run that starts at that pixel.
    const int lane_off = (ty + R) * lp + tx * P + R;
    // 1/var per kernel, after the window: read once per direction at the
    // selected scale (the staging barrier orders these writes)
    const int K = ks.D * ks.S;
    float *inv_lds = tile + lds_floats(TW + 1, TH, R, PM);
    for (int i = tid; i < K; i += NT) inv_lds[i] = as_const(tab)[2 * K + i];
    if constexpr (MODE == kWalkRing) {
        uint32_t *q = (uint32_t *)tile;
        if (stage_prefix<TW, TH, NT>(q, lp, x, W, H, pitch, x0, y0, R, rings.qscale, rings.qrange)) {
            walk_lane<P, kWalkRing, BODY>(q + lane_off, lp, ks, rings, tab, inv_lds, y, map, est, W, H, pitch, px0,
                                          py);
            return;
        }
        // A value of the window lies outside the fixed-point range (or is
        // not finite): float ring sums.  Every lane takes this branch.
        stage_tile<TW, TH, NT>(tile + 1, lp, x, W, H, pitch, x0, y0, R);
        __syncthreads();
        walk_lane<P, kWalkRingFloat, BODY>(tile + 1 + lane_off, lp, ks, rings, tab, inv_lds, y, map, est, W, H, pitch,
                                           px0, py);
    } else {
        // Pixels outside the image read as zero: the reference's zero padding.
        stage_tile<TW, TH, NT>(tile + 1, lp, x, W, H, pitch, x0, y0, R);
        __syncthreads();
        walk_lane<P, MODE, BODY>(tile + 1 + lane_off, lp, ks, rings, tab, inv_lds, y, map, est, W, H, pitch, px0,
                                 py);
    }
}

// Per-direction estimates at frozen scales (the scale maps a forward wrote):
// est[d][p] = the estimate of direction d at scale s_d(p), from the same walk
// and arithmetic as the forward with the same walk mode, so it is the value
// the forward aggregated, bit for bit.  Per lane, scales are walked up to the
// deepest one its pixels selected (ring forms: running sums; weighted walks:
// only the selected kernels).
template <int P, int MODE, int BODY, typename T>
__device__ __forceinline__ void estimate_lane(const T *__restrict__ src, int lp, const DevKset &ks,
                                              const DevRings &rings, const uint8_t *__restrict__ map,
                                              float *__restrict__ est, int W, int H, int px0, int py)
{
    constexpr bool EXACT = MODE == kWalkExact;
    const int S = ks.S;
    const size_t plane = (size_t)W * H;
    const size_t pix = (size_t)py * W + px0;
    bool inside[P];
#pragma unroll
    for (int p = 0; p < P; ++p) inside[p] = py < H && px0 + p < W;
    for (int d = 0; d < ks.D; ++d) {
        const auto gmap = as_global(map) + (size_t)d * plane + pix;
        int sd[P];
        float out[P], ring_sum[P];
        uint32_t ring_q[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            sd[p] = inside[p] ? min((int)gmap[p], S - 1) : S;
            out[p] = 0.0f;
            ring_sum[p] = 0.0f;
            ring_q[p] = 0u;
        }
        for (int s = 0; s < S; ++s) {
            bool need = false, here = false;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                need |= sd[p] >= s && sd[p] < S;
                here |= sd[p] == s;
            }
            if (!__any(need)) break;
            const int k = d * S + s;
            float acc[P];
            if constexpr (MODE == kWalkRing) {
                if (need) prefix_pairs<P>(src, rings.pairs, load_at(rings.pair_kern, (unsigned)k), ring_q);
                const float wq = load_at(rings.scale_q, (unsigned)k);
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = mul_rounded((float)(int)ring_q[p], wq);
            } else if constexpr (MODE == kWalkRingFloat) {
                if (need)
                    ring_sums<P>(src, rings.runs_b, load_at(rings.kern, (unsigned)k), ring_sum);
                const float w = as_const(rings.scale)[k];
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = ring_sum[p] * w;
            } else {
                if (!__any(here)) continue;
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = 0.0f;
                if (here) correlate<P, EXACT, BODY>(src, lp, ks, k, acc);
            }
#pragma unroll
            for (int p = 0; p < P; ++p)
                if (sd[p] == s) out[p] = acc[p];
        }
        const auto gest = as_global(est) + (size_t)d * plane + pix;
#pragma unroll
        for (int p = 0; p < P; ++p)
            if (inside[p]) gest[p] = out[p];
    }
}

// Persistent estimates kernel over a tile queue spanning a batch, on tile
// configuration SH's shape.  For the fixed-point ring walk the shape is the
// forward's own: the choice between fixed point and float ring sums is made
// per tile window, so the same windows give the same choice.
template <int SH, int MODE>
__global__ __launch_bounds__(kShapes[SH].nt) void estimate_kernel(DevKset ks, DevRings rings,
                                                                  const EstDesc *__restrict__ imgs,
                                                                  const uint2 *__restrict__ tiles, int num_tiles,
                                                                  int *__restrict__ counter, int R)
{
    extern __shared__ float tile[];
    __shared__ int next;
    constexpr TileShape sh = kShapes[SH];
    const int lp = prefix_pitch(sh.tw, R, sh.pm);  // the ring pairs' offsets are in this pitch
    const int tid = threadIdx.x;
    int tx, ty;
    lane_cell<sh.tw, sh.th, sh.px, sh.nt, sh.split>(tid, tx, ty);
    const int lane_off = (ty + R) * lp + tx * sh.px + R;
    for (;;) {
        if (tid == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 e = load_const(tiles + t);
        const EstDesc ed = load_const_struct(imgs + e.x);
        const ImgDesc &img = ed.img;
        const int x0 = (int)(e.y & 0xffffu) * sh.tw, y0 = (int)(e.y >> 16) * sh.th;
        const int px0 = x0 + tx * sh.px, py = y0 + ty;
        bool fixed = false;
        if constexpr (MODE == kWalkRing)
            fixed = stage_prefix<sh.tw, sh.th, sh.nt>((uint32_t *)tile, lp, img.x, img.W, img.H, img.pitch, x0, y0,
                                                      R, rings.qscale, rings.qrange);
        if (fixed) {
            estimate_lane<sh.px, kWalkRing, sh.body>((const uint32_t *)tile + lane_off, lp, ks, rings, img.map,
                                                      ed.est, img.W, img.H, px0, py);
        } else {
            __syncthreads();  // every lane has voted before the window is restaged
            stage_tile<sh.tw, sh.th, sh.nt>(tile + 1, lp, img.x, img.W, img.H, img.pitch, x0, y0, R);
            __syncthreads();
            constexpr int M = MODE == kWalkRing ? kWalkRingFloat : MODE;
            estimate_lane<sh.px, M, sh.body>(tile + 1 + lane_off, lp, ks, rings, img.map, ed.est, img.W, img.H, px0,
                                              py);
        }
        __syncthreads();  // the window and `next` are reused by the next tile
    }
}

// Plain form: one launch per image.  Work-groups are dealt to the 8 XCDs in
// turn, so block b runs tile (b % 8) * per_xcd + b / 8: consecutive tiles of
// one XCD are neighbours and share their halos in that XCD's L2.
template <int SH, int MODE>
__global__ __launch_bounds__(kShapes[SH].nt) __attribute__((amdgpu_waves_per_eu(kShapes[SH].wpe)))
void fwd_plain_kernel(DevKset ks, DevRings rings, ImgDesc img, int R, int num_tiles, int per_xcd)
{
    extern __shared__ float tile[];
    constexpr TileShape sh = kShapes[SH];
    const int b = blockIdx.x;
    const int t = (b & 7) * per_xcd + (b >> 3);
    if (t >= num_tiles) return;  // whole work-group, before any barrier
    const int tx = t % img.tiles_x, ty = t / img.tiles_x;
    process_tile<sh.tw, sh.th, sh.px, MODE, sh.body, sh.pm, sh.nt, sh.split>(
        tile, ks, rings, img.tab, img.x, img.y, img.map, img.est, img.W, img.H, img.pitch, tx * sh.tw, ty * sh.th, R);
}

// Batched form: a persistent grid drains one tile queue that spans every
// image of the batch.  tiles[t] = {image, (tile_y << 16) | tile_x}.  Every
// work-group leaves the loop once the counter passes num_tiles.
template <int SH, int MODE>
__global__ __launch_bounds__(kShapes[SH].nt) __attribute__((amdgpu_waves_per_eu(kShapes[SH].wpe)))
void fwd_batched_kernel(DevKset ks, DevRings rings, const ImgDesc *__restrict__ imgs,
                        const uint2 *__restrict__ tiles, int num_tiles, int *__restrict__ counter, int R)
{
    extern __shared__ float tile[];
    __shared__ int next;
    constexpr TileShape sh = kShapes[SH];
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 e = load_const(tiles + t);
        const ImgDesc img = load_const_struct(imgs + e.x);
        process_tile<sh.tw, sh.th, sh.px, MODE, sh.body, sh.pm, sh.nt, sh.split>(
            tile, ks, rings, img.tab, img.x, img.y, img.map, img.est, img.W, img.H, img.pitch,
            (int)(e.y & 0xffffu) * sh.tw,
            (int)(e.y >> 16) * sh.th, R);
        __syncthreads();  // the LDS tile and `next` are reused
    }
}

__global__ void reset_counter_kernel(int *counter) { *counter = 0; }

// LDS of one work-group: the window (with its leading column) and the
// [K] table of 1/var.
size_t fwd_lds_bytes(const TileShape &sh, int R, int K)
{
    return sizeof(float) * ((size_t)lds_floats(sh.tw + 1, sh.th, R, sh.pm) + (size_t)K);
}

// Launches that need more LDS than the default limit raise it first.
template <typename F>
hipError_t allow_lds(F *kernel, size_t lds)
{
    if (lds <= 65536) return hipSuccess;
    if (lds > (size_t)kMaxLdsBytes) return hipErrorInvalidValue;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds);
}

template <int SH, int MODE>
hipError_t plain_launch(const DevKset &ks, const DevRings &rings, const ImgDesc &img, int R, hipStream_t stream)
{
    constexpr TileShape sh = kShapes[SH];
    const int tiles_y = (img.H + sh.th - 1) / sh.th;
    const int num_tiles = img.tiles_x * tiles_y;
    const int per_xcd = (num_tiles + 7) / 8;
    const size_t lds = fwd_lds_bytes(sh, R, ks.D * ks.S);
    const hipError_t e = allow_lds(fwd_plain_kernel<SH, MODE>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fwd_plain_kernel<SH, MODE>), dim3(8 * per_xcd), dim3(sh.nt), lds, stream, ks, rings, img, R,
                       num_tiles, per_xcd);
    return hipGetLastError();
}

template <int SH>
hipError_t plain_dispatch(WalkMode mode, const DevKset &ks, const DevRings &rings, const ImgDesc &img, int R,
                          hipStream_t stream)
{
    switch (mode) {
    case kWalkExact: return plain_launch<SH, kWalkExact>(ks, rings, img, R, stream);
    case kWalkRing: return plain_launch<SH, kWalkRing>(ks, rings, img, R, stream);
    case kWalkRingFloat: return plain_launch<SH, kWalkRingFloat>(ks, rings, img, R, stream);
    default: return plain_launch<SH, kWalkFma>(ks, rings, img, R, stream);
    }
}

template <int SH, int MODE>
hipError_t batched_launch(const DevKset &ks, const DevRings &rings, const ImgDesc *imgs, const uint2 *tiles,
                          int num_tiles, int *counter, int R, int grid, hipStream_t stream)
{
    constexpr TileShape sh = kShapes[SH];
    const size_t lds = fwd_lds_bytes(sh, R, ks.D * ks.S);
    const hipError_t e = allow_lds(fwd_batched_kernel<SH, MODE>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fwd_batched_kernel<SH, MODE>), dim3(grid), dim3(sh.nt), lds, stream, ks, rings, imgs, tiles,
                       num_tiles, counter, R);
    return hipGetLastError();
}

template <int SH>
hipError_t batched_dispatch(WalkMode mode, const DevKset &ks, const DevRings &rings, const ImgDesc *imgs,
                            const uint2 *tiles, int num_tiles, int *counter, int R, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
    switch (mode) {
    case kWalkExact: return batched_launch<SH, kWalkExact>(ks, rings, imgs, tiles, num_tiles, counter, R, grid, stream);
    case kWalkRing: return batched_launch<SH, kWalkRing>(ks, rings, imgs, tiles, num_tiles, counter, R, grid, stream);
    case kWalkRingFloat:
        return batched_launch<SH, kWalkRingFloat>(ks, rings, imgs, tiles, num_tiles, counter, R, grid, stream);
    default: return batched_launch<SH, kWalkFma>(ks, rings, imgs, tiles, num_tiles, counter, R, grid, stream);
    }
}

template <int SH>
int occupancy(int R, int K)
{
    const size_t lds = fwd_lds_bytes(kShapes[SH], R, K);
    if (allow_lds(fwd_batched_kernel<SH, kWalkRing>, lds) != hipSuccess) return 1;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fwd_batched_kernel<SH, kWalkRing>, kShapes[SH].nt, lds) !=
        hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace

hipError_t launch_forward_plain(int shape, WalkMode mode, const DevKset &ks, const DevRings &rings,
                                const ImgDesc &img, int radius, hipStream_t stream)
{
    switch (shape) {
    case 0: return plain_dispatch<0>(mode, ks, rings, img, radius, stream);
    case 1: return plain_dispatch<1>(mode, ks, rings, img, radius, stream);
    case 2: return plain_dispatch<2>(mode, ks, rings, img, radius, stream);
    case 3: return plain_dispatch<3>(mode, ks, rings, img, radius, stream);
    case 4: return plain_dispatch<4>(mode, ks, rings, img, radius, stream);
    case 5: return plain_dispatch<5>(mode, ks, rings, img, radius, stream);
    case 6: return plain_dispatch<6>(mode, ks, rings, img, radius, stream);
    case 7: return plain_dispatch<7>(mode, ks, rings, img, radius, stream);
    case 8: return plain_dispatch<8>(mode, ks, rings, img, radius, stream);
    case 9: return plain_dispatch<9>(mode, ks, rings, img, radius, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_forward_batched(int shape, WalkMode mode, const DevKset &ks, const DevRings &rings,
                                  const ImgDesc *d_imgs, const uint2 *d_tiles, int num_tiles, int *d_counter,
                                  int radius, int grid, hipStream_t stream)
{
    switch (shape) {
#define LPA_B(SH)                                                                                                  \
    case SH: return batched_dispatch<SH>(mode, ks, rings, d_imgs, d_tiles, num_tiles, d_counter, radius, grid, stream);
        LPA_B(0) LPA_B(1) LPA_B(2) LPA_B(3) LPA_B(4) LPA_B(5) LPA_B(6) LPA_B(7) LPA_B(8) LPA_B(9)
#undef LPA_B
    default: return hipErrorInvalidValue;
    }
}

template <int SH, int MODE>
hipError_t estimate_launch(const DevKset &ks, const DevRings &rings, const EstDesc *d_imgs, const uint2 *d_tiles,
                           int num_tiles, int *d_counter, int radius, int grid, hipStream_t stream)
{
    constexpr TileShape sh = kShapes[SH];
    // two words past the window: the chain body reads one value past a run
    const size_t lds = sizeof(float) * ((size_t)lds_floats(sh.tw + 1, sh.th, radius, sh.pm) + 2);
    const hipError_t e = allow_lds(estimate_kernel<SH, MODE>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((estimate_kernel<SH, MODE>), dim3(grid), dim3(sh.nt), lds, stream, ks, rings, d_imgs, d_tiles,
                       num_tiles, d_counter, radius);
    return hipGetLastError();
}

hipError_t launch_estimates(WalkMode mode, int shape, const DevKset &ks, const DevRings &rings,
                            const EstDesc *d_imgs, const uint2 *d_tiles, int num_tiles, int *d_counter, int radius,
                            int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
    if (mode == kWalkRing) {
        switch (shape) {
#define LPA_E(SH)                                                                                                  \
    case SH: return estimate_launch<SH, kWalkRing>(ks, rings, d_imgs, d_tiles, num_tiles, d_counter, radius, grid, stream);
            LPA_E(0) LPA_E(1) LPA_E(2) LPA_E(3) LPA_E(4) LPA_E(5) LPA_E(6) LPA_E(7) LPA_E(8) LPA_E(9)
#undef LPA_E
        default: return hipErrorInvalidValue;
        }
    }
    if (shape != kEstimateShape) return hipErrorInvalidValue;
    constexpr int SH = kEstimateShape;
    switch (mode) {
    case kWalkFma: return estimate_launch<SH, kWalkFma>(ks, rings, d_imgs, d_tiles, num_tiles, d_counter, radius, grid, stream);
    case kWalkExact:
        return estimate_launch<SH, kWalkExact>(ks, rings, d_imgs, d_tiles, num_tiles, d_counter, radius, grid, stream);
    case kWalkRingFloat:
        return estimate_launch<SH, kWalkRingFloat>(ks, rings, d_imgs, d_tiles, num_tiles, d_counter, radius, grid,
                                                   stream);
    default: return hipErrorInvalidValue;
    }
}

int blocks_per_cu(int shape, int radius, int num_kernels)
{
    switch (shape) {
    case 0: return occupancy<0>(radius, num_kernels);
    case 1: return occupancy<1>(radius, num_kernels);
    case 2: return occupancy<2>(radius, num_kernels);
    case 3: return occupancy<3>(radius, num_kernels);
    case 4: return occupancy<4>(radius, num_kernels);
    case 5: return occupancy<5>(radius, num_kernels);
    case 6: return occupancy<6>(radius, num_kernels);
    case 7: return occupancy<7>(radius, num_kernels);
    case 8: return occupancy<8>(radius, num_kernels);
    case 9: return occupancy<9>(radius, num_kernels);
    default: return 1;
    }
}

}  // namespace lpa
