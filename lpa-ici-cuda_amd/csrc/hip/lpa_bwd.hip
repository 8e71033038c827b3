// Adjoint of the LPA-ICI forward with the selected scales frozen.
//
// With s_d(p) the scale the forward selected for pixel p and direction d,
// k = d*S + s_d(p), inv_d(p) = 1/var_k, Wt(p) = sum_d inv_d(p),
// yhat_d(p) = sum_t w_k[t] x[p+t] and y(p) = sum_d inv_d yhat_d / Wt, the
// output is linear in x and rational in the taps (var_k = sigma^2 sum w_k^2).
// For an upstream gradient g, with q(p) = g(p)/Wt(p) and c_d(p) = q inv_d:
//   dL/dx[u]     = sum_d sum_p c_d(p) w_{d,s_d(p)}[u - p]
//   dL/dw_k[t]   = A_k[t] - 2 w_k[t] E_k / sumsq_k, where
//   A_k[t]       = sum_{p: k(p,d)=k} c_d(p) x[p+t]
//   E_k          = sum_{p: k(p,d)=k} c_d(p) (yhat_d(p) - y(p)).
// Nothing of the forward but y and the uint8 scale maps is kept: yhat is
// recomputed from the maps here.
//
// Pass 1 (one tile at a time): q, the recomputed yhat of the selected scale
// and E, then A per wave and selected kernel (a walk over the kernel's runs
// for the whole wave, or the kernel's taps on the lanes for a few pixels),
// added into the work-group's row of `partial`.
// Pass 2 (dx): per direction and scale, the masked map q*inv*[s_d = s] over
// tile + halo is correlated with the 180-degree rotated kernel (for ring-form
// sets, nested masked maps with the rotated rings).
// Pass 3: rows of `partial` are summed and the variance term is added.
#include <hip/hip_runtime.h>

#include "lpa_device.hpp"
#include "lpa_tile.hpp"

namespace lpa {
namespace {

constexpr int kMaxScales = 16;

// Pass-1 tile: 64 x 64 output pixels, 8 adjacent pixels per lane, 512 lanes
// held to 128 registers (two work-groups per CU).
constexpr int GW = kBwdTileW, GH = kBwdTileH, GP = 8, GLX = GW / GP, GNT = GLX * GH;
static_assert(GNT % 64 == 0 && GNT <= 1024, "pass-1 lanes: whole waves, at most 1024");
// A wave covers a 16 x 32 block of the tile (the 8 waves: 2 across, 4 down),
// which holds fewer distinct selected scales than a 8 x 64 strip, so fewer
// kernels are walked per wave.  Each half-wave (whose 32 lanes the LDS serves
// together) takes 8 rows of it: 4 lanes per row, 8 px apart, on rows 9 (float
// tile) or 1 (prefix table) banks apart, so the 32 lanes of a read hit 32
// distinct banks.  Columns are in units of GP pixels.
static_assert(GLX == 8 && GNT == 512, "the lane layout assumes 8 waves on a 64 x 64 tile");
__device__ __forceinline__ int grad_lane_col(int lane) { return lane & 3; }
__device__ __forceinline__ int grad_lane_row(int lane) { return (lane >> 5) * 8 + ((lane & 31) >> 2); }
__device__ __forceinline__ int grad_wave_col(int wave) { return (wave & 1) * 4; }
__device__ __forceinline__ int grad_wave_row(int wave) { return (wave >> 1) * 16; }

// The float tile (pitch lds_pitch(GW, R, kBwdPitchResidue)) and the prefix
// table (one more column, pitch lds_pitch(GW + 1, R)) share the start of the
// LDS; the [3][K] table follows the larger of the two.
__host__ __device__ inline int grad_tab_offset(int R)
{
    const int a = lds_floats(GW, GH, R, kBwdPitchResidue), b = lds_floats(GW + 1, GH, R);
    return a > b ? a : b;
}

__global__ __launch_bounds__(kThreads) void fill_kernel(float *p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void reset_counter2(int *c)
{
    c[0] = 0;
    c[1] = 0;
}

// Cross-lane helpers.  swap32 / swap16 exchange two registers across lane
// bit 5 / lane bit 4 (v_permlane32_swap / v_permlane16_swap): afterwards
// a + b holds, in every lane, its own half plus its partner's copy of the
// same half.  row_total sums the 16 lanes of each row with four DPP adds.
__device__ __forceinline__ void swap32(float &a, float &b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

__device__ __forceinline__ void swap16(float &a, float &b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float row_total(float v)
{
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp<0x124>(v);  // row_ror:4
    v += dpp<0x128>(v);  // row_ror:8
    return v;
}

// Sum of v over the 64 lanes, in every lane.
__device__ __forceinline__ float wave_total(float v)
{
    float a = v, b = v;
    swap32(a, b);
    float c = a + b, d = c;
    swap16(c, d);
    return row_total(c + d);
}

// Bit s: some pixel of the lane selected scale s (scale indices of pixels
// outside the image are 255 and set no bit).
__device__ __forceinline__ unsigned lane_scales(const int (&sd)[GP])
{
    unsigned b = 0;
#pragma unroll
    for (int p = 0; p < GP; ++p) b |= sd[p] < 32 ? 1u << sd[p] : 0u;
    return b;
}

// OR of v over the 64 lanes of a wave, wave-uniform.
__device__ __forceinline__ unsigned wave_or(unsigned v)
{
    float a = __uint_as_float(v), b = a;
    swap32(a, b);
    float c = __uint_as_float(__float_as_uint(a) | __float_as_uint(b)), d = c;
    swap16(c, d);
    unsigned u = __float_as_uint(c) | __float_as_uint(d);
    u |= __float_as_uint(dpp<0xB1>(__uint_as_float(u)));   // quad_perm [1,0,3,2]
    u |= __float_as_uint(dpp<0x4E>(__uint_as_float(u)));   // quad_perm [2,3,0,1]
    u |= __float_as_uint(dpp<0x124>(__uint_as_float(u)));  // row_ror:4
    u |= __float_as_uint(dpp<0x128>(__uint_as_float(u)));  // row_ror:8
    return __builtin_amdgcn_readfirstlane(u);
}

// Sum L (<= 8) values over the 64 lanes of a wave and add the first n (<= L)
// totals into row[0..n) with fire-and-forget float atomics.  Reduce-scatter: the swaps
// halve the values per lane across lane bits 5 and 4 without any select, so
// each lane keeps Lp/4 partial sums (Lp = L rounded up to a power of two),
// which the row reductions finish; lanes 0, 16, 32 and 48 own
// j = b5 * Lp/2 + b4 * Lp/4 + i.
template <int L>
__device__ __forceinline__ void reduce_scatter_add(const float (&v)[L], float *__restrict__ row, int lane, int n = L)
{
    static_assert(L >= 1 && L <= 8, "pieces of at most 8 taps");
    constexpr int Lp = L <= 1 ? 1 : L <= 2 ? 2 : L <= 4 ? 4 : 8;
    if constexpr (Lp == 1) {
        const float t = wave_total(v[0]);
        if (lane == 0 && n > 0) unsafeAtomicAdd(row, t);
    } else if constexpr (Lp == 2) {
        float a = v[0], b = L > 1 ? v[1] : 0.0f;
        swap32(a, b);
        float c = a + b, d = c;
        swap16(c, d);
        const float t = row_total(c + d);
        const int j = lane >> 5;
        if ((lane & 31) == 0 && j < n) unsafeAtomicAdd(row + j, t);
    } else {
        float h[Lp / 2];
#pragma unroll
        for (int i = 0; i < Lp / 2; ++i) {
            float a = v[i], b = i + Lp / 2 < L ? v[i + Lp / 2] : 0.0f;
            swap32(a, b);
            h[i] = a + b;
        }
#pragma unroll
        for (int i = 0; i < Lp / 4; ++i) {
            float a = h[i], b = h[i + Lp / 4];
            swap16(a, b);
            const float t = row_total(a + b);
            const int j = (lane >> 5) * (Lp / 2) + ((lane >> 4) & 1) * (Lp / 4) + i;
            if ((lane & 15) == 0 && j < n) unsafeAtomicAdd(row + j, t);
        }
    }
}

// One run piece of L taps for the lane's GP pixels: yacc[p] += sum_j w[j] x,
// and the wave's A partials for the piece's taps, sum_p m[p] x[p + j], added
// into the wave's row at the piece's first run weight.
template <int L, bool YHAT>
__device__ __forceinline__ void grad_piece(const float *__restrict__ src, const float *__restrict__ w,
                                           const float (&m)[GP], float (&yacc)[GP], float *__restrict__ row, int lane)
{
    float v[L + GP - 1];
#pragma unroll
    for (int q = 0; q < L + GP - 1; ++q) v[q] = src[q];
    if constexpr (YHAT) {
#pragma unroll
        for (int j = 0; j < L; ++j) {
            const float c = as_const(w)[j];
#pragma unroll
            for (int p = 0; p < GP; ++p) yacc[p] = __builtin_fmaf(c, v[j + p], yacc[p]);
        }
    }
    float part[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        float a = m[0] * v[j];
#pragma unroll
        for (int p = 1; p < GP; ++p) a = __builtin_fmaf(m[p], v[j + p], a);
        part[j] = a;
    }
    reduce_scatter_add<L>(part, row, lane);
}

// The lane's GP scale indices in one direction's map (255 for pixels outside
// the image): one 8-byte load where the lane's pixels are whole and aligned
// (`whole`: an 8-aligned map of an image whose width is a multiple of 8),
// else byte by byte.
__device__ __forceinline__ void load_scales(const uint8_t *__restrict__ m, bool whole, const bool (&inside)[GP],
                                            int (&sd)[GP])
{
    if (whole) {
        const unsigned long long v =
            *reinterpret_cast<const __attribute__((address_space(1))) unsigned long long *>(as_global(m));
#pragma unroll
        for (int p = 0; p < GP; ++p) sd[p] = (int)((v >> (8 * p)) & 255u);
    } else {
        const auto g = as_global(m);
#pragma unroll
        for (int p = 0; p < GP; ++p) sd[p] = inside[p] ? g[p] : 255;
    }
}

// A partials of one kernel for the few pixels of a wave that selected it,
// with the kernel's taps on the lanes (tap t = 64 i + lane, IT = number of
// 64-tap slices): every pixel p with coefficient c_p adds c_p x[p + t] to its
// lane's sums, then each lane adds its taps' sums into the row.  Cost grows
// with the pixels, not with the wave: one LDS read and one FMA per pixel and
// tap, against the tile walk's sliding window over all 512 pixels of the wave.
// The tap offsets in the tile are read from tap_lds, one load per slice.
// `pix_off` is the tile offset of the lane's first pixel from base0
// (grad_lane_row * lp + grad_lane_col * GP).  Offsets are taken in bytes,
// the lanes' pixel offsets and the tap offsets once per call, so a pixel's
// address is one readlane and each read one add.
template <int IT>
__device__ __forceinline__ void taps_on_lanes(const float *__restrict__ xt, int base0, int pix_off,
                                              const int *__restrict__ tap_lds, int w0, int taps, const float (&m)[GP],
                                              const int (&sd)[GP], int s, float *__restrict__ row, int lane)
{
    // the lane's tap addresses, relative to the pixel offset added per read
    const char *xo[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int t = 64 * i + lane;
        xo[i] = reinterpret_cast<const char *>(xt + (t < taps ? as_global(tap_lds)[w0 + t] : 0));
    }
    const int pb = 4 * (base0 + pix_off);
    float acc[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) acc[i] = 0.0f;
    // lane L holds pixels grad_lane_col(L) * GP + p of row grad_lane_row(L) of the wave's block
    auto pixel = [&](int L, int p, float &c) {
        c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m[p]), L));
        return __builtin_amdgcn_readlane(pb, L) + 4 * p;
    };
    auto at = [&](int b, int i) { return *reinterpret_cast<const float *>(xo[i] + b); };
#pragma unroll
    for (int p = 0; p < GP; ++p) {
        unsigned long long mask = __ballot(sd[p] == s);
        // two pixels per step while two are left, so their reads overlap
        for (int n = __builtin_popcountll(mask); n >= 2; n -= 2) {
            const int L0 = __builtin_ctzll(mask);
            mask ^= 1ull << L0;
            const int L1 = __builtin_ctzll(mask);
            mask ^= 1ull << L1;
            float c0, c1;
            const int b0 = pixel(L0, p, c0), b1 = pixel(L1, p, c1);
            float v0[IT], v1[IT];
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                v0[i] = at(b0, i);
                v1[i] = at(b1, i);
            }
#pragma unroll
            for (int i = 0; i < IT; ++i) acc[i] = __builtin_fmaf(c1, v1[i], __builtin_fmaf(c0, v0[i], acc[i]));
        }
        if (mask) {
            float c0;
            const int b0 = pixel(__builtin_ctzll(mask), p, c0);
#pragma unroll
            for (int i = 0; i < IT; ++i) acc[i] = __builtin_fmaf(c0, at(b0, i), acc[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int t = 64 * i + lane;
        if (t < taps) unsafeAtomicAdd(row + w0 + t, acc[i]);
    }
}

// The cheaper of the two A walks for one kernel and wave, in measured time
// units: the tile walk costs about kPieceCost per run piece whatever the
// pixels; taps on lanes about kLanePixelCost per pixel and 64-tap slice (the
// pass time is flat for lane costs from 6 to 12 on the shipped set).
constexpr int kPieceCost = 100, kLanePixelCost = 8, kMaxLaneSlices = 4;

// Pass 1.  For every tile: q = g / sum_d inv_d (written to the q map for pass
// 2), then two phases.
//   yhat and E (RING): yhat of the selected scale is recomputed with the
//   forward's own arithmetic, so it equals the estimate y was built from bit
//   for bit (E is a difference yhat - y and enters dL/dw beside A, where the
//   two can cancel): the ring sums up to the wave's deepest selected scale, in
//   fixed point from row-prefix tables when every value of this window is in
//   range (the forward decides the same per window of its own tile, so the
//   two agree unless an out-of-range value lies within reach of one window and
//   not the other), else float ring sums over the float tile.  Then E per
//   scale from the wave.
//   A: the float tile; per direction and for every scale selected by some
//   pixel of the wave, the A partials of that kernel with the masked
//   coefficients m_p = [s_d(p) = s] q_p inv, either by one walk over the
//   kernel's runs for the wave's 512 pixels (partials summed across the lanes
//   by a reduce-scatter) or, when few pixels of the wave selected the kernel,
//   with its taps on the lanes (taps_on_lanes).  Without a ring form the walk
//   also accumulates yhat, and E follows it.
// Each work-group owns one row of `partial` ([grid][NW + K]); its waves add
// into it with float atomics that return nothing.
template <bool RING>
__global__ __launch_bounds__(GNT) __attribute__((amdgpu_waves_per_eu(GNT > kThreads ? 4 : 1)))
void bwd_grad_kernel(DevKset ks, DevRings rings, DevAdjoint adj,
                                                            const BwdDesc *__restrict__ imgs,
                                                            const uint2 *__restrict__ tiles, int num_tiles,
                                                            int *__restrict__ counter, int R,
                                                            float *__restrict__ partial)
{
    extern __shared__ __align__(16) float lds_f[];
    __shared__ int next;
    const int S = ks.S, D = ks.D, K = D * S;
    const int lp = lds_pitch(GW, R, kBwdPitchResidue), lpq = prefix_pitch(GW, R);  // lpq: the ring pairs' pitch
    float *xt = lds_f;
    uint32_t *qt = (uint32_t *)lds_f;
    float *tab = lds_f + grad_tab_offset(R);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int NW = adj.num_weights;
    float *__restrict__ row = partial + (size_t)blockIdx.x * (NW + K);

    for (;;) {
        if (tid == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 te = load_const(tiles + t);
        const BwdDesc bd = load_const_struct(imgs + te.x);
        const ImgDesc &img = bd.img;
        const int x0 = (int)(te.y & 0xffffu) * GW, y0 = (int)(te.y >> 16) * GH;
        for (int i = tid; i < 3 * K; i += GNT) tab[i] = as_const(img.tab)[i];
        // fixed: the forward's fixed-point ring sums apply to this window
        // (every value within the range); the vote is also the barrier.
        // With the forward's estimates at hand nothing is recomputed, and the
        // float tile is staged at once.
        const bool have_est = img.est != nullptr;
        bool fixed = false;
        if constexpr (RING)
            if (!have_est)
                fixed = stage_prefix<GW, GH, GNT>(qt, lpq, img.x, img.W, img.H, img.pitch, x0, y0, R, rings.qscale,
                                                  rings.qrange);
        if (!fixed) {
            __syncthreads();  // every lane has voted before the window is restaged
            stage_tile<GW, GH, GNT>(xt, lp, img.x, img.W, img.H, img.pitch, x0, y0, R);
            __syncthreads();
        }

        const float *inv_t = tab + 2 * K;
        const int tx = grad_wave_col(wave) + grad_lane_col(lane), ty = grad_wave_row(wave) + grad_lane_row(lane);
        const int px0 = x0 + tx * GP, py = y0 + ty;
        const float *lane_x = xt + (ty + R) * lp + tx * GP + R;
        const size_t plane = (size_t)img.W * img.H;
        const size_t pix = (size_t)py * img.W + px0;
        const auto gg = as_global(bd.g) + (size_t)py * img.pitch + px0;
        const auto gy = as_global(img.y) + (size_t)py * img.pitch + px0;
        const auto gq = as_global(bd.q) + pix;
        bool inside[GP];
        float q[GP], yv[GP], wt[GP];
#pragma unroll
        for (int p = 0; p < GP; ++p) {
            inside[p] = py < img.H && px0 + p < img.W;
            wt[p] = 0.0f;
        }
        const bool whole = py < img.H && px0 < img.W && img.W % GP == 0 &&
                           ((uintptr_t)img.map & (GP - 1)) == 0;
        for (int d = 0; d < D; ++d) {
            int sd[GP];
            load_scales(img.map + (size_t)d * plane + pix, whole, inside, sd);
#pragma unroll
            for (int p = 0; p < GP; ++p)
                if (inside[p]) wt[p] += inv_t[d * S + min(sd[p], S - 1)];
        }
#pragma unroll
        for (int p = 0; p < GP; ++p) {
            q[p] = inside[p] ? gg[p] / wt[p] : 0.0f;
            yv[p] = inside[p] ? gy[p] : 0.0f;
            if (inside[p]) gq[p] = q[p];
        }

        if constexpr (RING) {
            // The prefix value Q[.][R + tx*GP] is the left boundary of a run
            // that starts at the lane's first pixel.
            const uint32_t *lane_q = qt + (ty + R) * lpq + tx * GP + R;
            for (int d = 0; d < D; ++d) {
                int sd[GP];
                load_scales(img.map + (size_t)d * plane + pix, whole, inside, sd);
                // scales selected by the lane's pixels, and by the wave's
                const unsigned ls = lane_scales(sd), ws = wave_or(ls);
                float yh[GP];
                if (have_est) {
                    const auto ge = as_global(img.est) + (size_t)d * plane + pix;
#pragma unroll
                    for (int p = 0; p < GP; ++p) yh[p] = inside[p] ? ge[p] : 0.0f;
                } else {
                    // The forward's ring walk: running sum over rings 0..s up to
                    // the deepest selected scale; yhat = (w_s 2^-F) * (fixed-point
                    // sum), or w_s * (float sum) for a window outside the range.
                    float rs[GP];
                    uint32_t rq[GP];
#pragma unroll
                    for (int p = 0; p < GP; ++p) {
                        yh[p] = 0.0f;
                        rs[p] = 0.0f;
                        rq[p] = 0u;
                    }
                    for (int s = 0; s < S; ++s) {
                        if ((ws >> s) == 0u) break;  // no pixel of the wave selected s or deeper
                        const bool need = (ls >> s) != 0u;
                        const int k = d * S + s;
                        if (fixed) {
                            if (need) prefix_pairs<GP>(lane_q, rings.pairs, load_at(rings.pair_kern, (unsigned)k), rq);
                            const float wq = load_at(rings.scale_q, (unsigned)k);  // w_s 2^-F, as the forward
#pragma unroll
                            for (int p = 0; p < GP; ++p)
                                if (sd[p] == s) yh[p] = mul_rounded((float)(int)rq[p], wq);
                        } else {
                            if (need)
                                ring_sums<GP>(lane_x, rings.runs_b, load_at(rings.kern, (unsigned)k), rs);
                            const float w = as_const(rings.scale)[k];
#pragma unroll
                            for (int p = 0; p < GP; ++p)
                                if (sd[p] == s) yh[p] = rs[p] * w;
                        }
                    }
                }
                // E_k: sum over the pixels that selected s of c (yhat - y),
                // c = q inv_k.  Each lane sums its pixels per scale; the wave
                // then sums 8 scales at a time with one reduce-scatter.
                float dv[GP];
#pragma unroll
                for (int p = 0; p < GP; ++p) dv[p] = yh[p] - yv[p];
                for (int s0 = 0; s0 < S; s0 += 8) {
                    float e[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int s = s0 + j;
                        float a = 0.0f;
                        if ((ws >> s) & 1u) {
#pragma unroll
                            for (int p = 0; p < GP; ++p)
                                if (sd[p] == s) a = __builtin_fmaf(q[p], dv[p], a);
                            a *= inv_t[d * S + s];
                        }
                        e[j] = a;
                    }
                    reduce_scatter_add<8>(e, row + NW + d * S + s0, lane, min(8, S - s0));
                }
            }
            if (fixed) {
                __syncthreads();  // the prefix table is done with
                stage_tile<GW, GH, GNT>(xt, lp, img.x, img.W, img.H, img.pitch, x0, y0, R);
                __syncthreads();
            }
        }

        for (int d = 0; d < D; ++d) {
            int sd[GP];
            load_scales(img.map + (size_t)d * plane + pix, whole, inside, sd);
            const unsigned ws = wave_or(lane_scales(sd));  // scales selected in the wave
            for (int s = 0; s < S; ++s) {
                if (!((ws >> s) & 1u)) continue;
                const int k = d * S + s;
                const float inv = inv_t[k];
                float m[GP], yacc[GP];
#pragma unroll
                for (int p = 0; p < GP; ++p) {
                    m[p] = sd[p] == s ? q[p] * inv : 0.0f;
                    yacc[p] = 0.0f;
                }
                const int2 kr = load_const(ks.kern + k);
                if constexpr (RING) {
                    // Few pixels of the wave selected this kernel: taps on lanes.
                    int n = 0;
#pragma unroll
                    for (int p = 0; p < GP; ++p) n += __popcll(__ballot(sd[p] == s));
                    const int4 kw = load_const(adj.kern_w + k);
                    const int slices = (kw.y + 63) >> 6;
                    if (slices <= kMaxLaneSlices && n * slices * kLanePixelCost < kw.z * kPieceCost) {
                        const int base0 = (grad_wave_row(wave) + R) * lp + grad_wave_col(wave) * GP + R;
                        const int po = grad_lane_row(lane) * lp + grad_lane_col(lane) * GP;
                        switch (slices) {
                        case 1: taps_on_lanes<1>(xt, base0, po, adj.tap_lds, kw.x, kw.y, m, sd, s, row, lane); break;
                        case 2: taps_on_lanes<2>(xt, base0, po, adj.tap_lds, kw.x, kw.y, m, sd, s, row, lane); break;
                        case 3: taps_on_lanes<3>(xt, base0, po, adj.tap_lds, kw.x, kw.y, m, sd, s, row, lane); break;
                        default: taps_on_lanes<4>(xt, base0, po, adj.tap_lds, kw.x, kw.y, m, sd, s, row, lane); break;
                        }
                        continue;
                    }
                }
                for (int r = 0; r < kr.y; ++r) {
                    const int4 run = load_const(ks.runs + kr.x + r);
                    for (int j0 = 0; j0 < run.z; j0 += kGradPiece) {
                        const int off = __builtin_amdgcn_readfirstlane(run.x * lp + run.y + j0);
                        const float *src = lane_x + off;
                        const float *__restrict__ w = ks.weights + run.w + j0;
                        float *__restrict__ rrow = row + run.w + j0;
                        switch (min(kGradPiece, run.z - j0)) {
#define LPA_GRAD_CASE(L)                                                                                           \
    case L: grad_piece<L, !RING>(src, w, m, yacc, rrow, lane); break;
                            LPA_GRAD_CASE(1) LPA_GRAD_CASE(2) LPA_GRAD_CASE(3) LPA_GRAD_CASE(4) LPA_GRAD_CASE(5)
                            LPA_GRAD_CASE(6) LPA_GRAD_CASE(7) LPA_GRAD_CASE(8)
#undef LPA_GRAD_CASE
                        default: break;
                        }
                    }
                }
                if constexpr (!RING) {
                    // E_k from the walk's own yhat.
                    float e = 0.0f;
#pragma unroll
                    for (int p = 0; p < GP; ++p) e = __builtin_fmaf(m[p], yacc[p] - yv[p], e);
                    e = wave_total(e);
                    if (lane == 0) unsafeAtomicAdd(row + NW + k, e);
                }
            }
        }
        __syncthreads();  // the LDS tile, the table and `next` are reused by the next tile
    }
}

// Unweighted run sums for the P pixels starting at `src` (the masked map of
// the dx pass): acc[p] += sum_{j < L} src[p + j].  From L = 4 on, the first
// pixel's sum slides to the next pixels ((s + src[p + L - 1]) - src[p - 1]):
// L + 3P - 3 operations instead of L P.  The dx pass has no bit-exact
// counterpart to match (its bar is the FP64 adjoint); the two extra
// roundings per slide are far below it.
template <int L, int P>
__device__ __forceinline__ void run_sums_sliding(const float *__restrict__ src, float (&acc)[P])
{
    float v[L + P - 1];
#pragma unroll
    for (int q = 0; q < L + P - 1; ++q) v[q] = src[q];
    if constexpr (L < 4) {
#pragma unroll
        for (int j = 0; j < L; ++j)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] += v[j + p];
    } else {
        float sum = v[0];
#pragma unroll
        for (int j = 1; j < L; ++j) sum += v[j];
        acc[0] += sum;
#pragma unroll
        for (int p = 1; p < P; ++p) {
            sum = (sum + v[p + L - 1]) - v[p - 1];
            acc[p] += sum;
        }
    }
}

// Ring sums of the masked map for the P pixels of a lane.  The switch has a
// case per run length up to 16; a longer run (up to 32) takes its first 16
// taps as a run of its own, which keeps the register window at 16 + P - 1.
constexpr int kDxMaxCase = 16;
static_assert(kMaxRun <= 2 * kDxMaxCase, "one split covers every run");
template <int P>
__device__ __forceinline__ void ring_sums_sliding(const float *__restrict__ lane, int lp, const int4 *__restrict__ runs,
                                                  int2 kr, float (&acc)[P])
{
    for (int r = 0; r < kr.y; ++r) {
        const int4 run = load_const(runs + kr.x + r);
        const float *__restrict__ src = lane + __builtin_amdgcn_readfirstlane(run.x * lp + run.y);
        int L = run.z;
        if (L > kDxMaxCase) {
            run_sums_sliding<kDxMaxCase, P>(src, acc);
            src += kDxMaxCase;
            L -= kDxMaxCase;
        }
        switch (L) {
#define LPA_SLIDE_CASE(L)                                                                                          \
    case L: run_sums_sliding<L, P>(src, acc); break;
            LPA_SLIDE_CASE(1) LPA_SLIDE_CASE(2) LPA_SLIDE_CASE(3) LPA_SLIDE_CASE(4) LPA_SLIDE_CASE(5)
            LPA_SLIDE_CASE(6) LPA_SLIDE_CASE(7) LPA_SLIDE_CASE(8) LPA_SLIDE_CASE(9) LPA_SLIDE_CASE(10)
            LPA_SLIDE_CASE(11) LPA_SLIDE_CASE(12) LPA_SLIDE_CASE(13) LPA_SLIDE_CASE(14) LPA_SLIDE_CASE(15)
            LPA_SLIDE_CASE(16)
#undef LPA_SLIDE_CASE
        default: break;
        }
    }
}

// Pass 2 (dx), on tiles 64 wide and 16 P tall, 1024 lanes of P pixels along
// a row: P = 8 (64 x 128 tiles) for kernel sets in ring form, P = 4 (64 x 64)
// for weighted sets, whose FMA window would not fit P = 8 in registers.  The
// scalar work of a ring run (its record, the switch) is paid once per lane
// whatever P, so the wider lane halves it per pixel.  Each lane keeps q and
// the scale index of the tile+halo elements (rows tid/64 + 16i, columns
// tid%64 + 64j) in registers; the LDS holds only the masked map.  A wave
// covers P rows: each half-wave a 32-pixel-wide half of them (32 / P lanes
// per row, P px apart, rows one bank apart: the 32 lanes of a read fall in
// 32 banks).
constexpr int DXW = 64, DXT = 1024;
constexpr int DXLP = lds_pitch(DXW, kMaxRadius);  // the map's row pitch, 1 mod 32
static_assert(DXLP % 32 == 1, "rows one bank apart");
template <bool RING> struct DxShape {
    static constexpr int P = RING ? 8 : 4;                   // pixels per lane, along a row
    static constexpr int H = 16 * P;                         // tile rows: 16 waves of P rows
    static constexpr int LH = 32 / P;                        // lanes per row in a half-wave
    static constexpr int SLOT_ROWS = (H + 2 * kMaxRadius + 15) / 16;  // rows r0 + 16 i cover H + 2R
    static constexpr int SLOT_COLS = 2;  // columns c0 + 64 j cover 64 + 2R
    static_assert(DXW + 2 * kMaxRadius <= 64 * SLOT_COLS, "the slots cover the widest window");
};
static_assert(DxShape<true>::H == kDxTileRowsRing && DxShape<false>::H == kDxTileRowsWeighted, "tile lists");

//
// RING (kernel sets in ring form, k_{d,s} = w_{d,s} on support(d,s), supports
// nested): with M_{d,r}(p) = q(p) inv_{d,s} w_{d,s} [s = s_d(p) >= r],
//   dL/dx[u] = sum_d sum_r sum_{t in ring r} M_{d,r}(u - t),
// since support(d,s) is the union of the rings r <= s.  Walking r from S-1
// down to 0 the map only gains the pixels with s_d(p) = r, and each
// 180-degree-rotated ring is applied once per direction: float adds over the
// ring taps instead of FMAs over every tap of every scale present.
template <bool RING>
__global__ __launch_bounds__(DXT) void bwd_dx_kernel(DevKset ks, DevRings rings, DevAdjoint adj,
                                                     const BwdDesc *__restrict__ imgs, const uint2 *__restrict__ tiles,
                                                     int num_tiles, int *__restrict__ counter, int R)
{
    using SH = DxShape<RING>;
    constexpr int P = SH::P, DXH = SH::H, LH = SH::LH;
    constexpr int SR = SH::SLOT_ROWS, SC = SH::SLOT_COLS;
    static_assert(SR * SC <= 32, "slot validity bits fit one word");
    // The map's row pitch is that of the largest radius, whatever R: a
    // compile-time pitch puts every slot's LDS offset in the instruction.
    constexpr int lp = DXLP;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next;
    const int S = ks.S, D = ks.D, K = D * S;
    float *mt = (float *)lds;
    float4 *mt4 = (float4 *)lds;
    const int tid = threadIdx.x;
    const int fw = DXW + 2 * R, fh = DXH + 2 * R;
    const int map4 = (fh * lp + 3) / 4;  // float4s covering the map (the LDS size is rounded up to match)
    const int r0 = tid >> 6, c0 = tid & 63;

    for (;;) {
        if (tid == 0) next = atomicAdd(counter + 1, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 te = load_const(tiles + t);
        const BwdDesc bd = load_const_struct(imgs + te.x);
        const ImgDesc &img = bd.img;
        const int x0 = (int)(te.y & 0xffffu) * DXW, y0 = (int)(te.y >> 16) * DXH;
        const float *__restrict__ inv_t = img.tab + 2 * K;  // wave-uniform scalar loads
        const size_t plane = (size_t)img.W * img.H;

        // Slot (i, j) of this lane is the window element at row r0 + 16 i,
        // column c0 + 64 j, i.e. image offset lane_off + 16 i W + 64 j; bit
        // i SC + j of `valid` says it lies in the window and in the image.
        const int gx0 = x0 - R + c0, gy0 = y0 - R + r0;
        const long long lane_off = (long long)gy0 * img.W + gx0;
        unsigned valid = 0;
#pragma unroll
        for (int i = 0; i < SR; ++i)
#pragma unroll
            for (int j = 0; j < SC; ++j) {
                const int r = r0 + 16 * i, c = c0 + 64 * j, gy = gy0 + 16 * i, gx = gx0 + 64 * j;
                const bool in = r < fh && c < fw && gy >= 0 && gy < img.H && gx >= 0 && gx < img.W;
                valid |= (unsigned)in << (i * SC + j);
            }
        float qv[SR][SC];
#pragma unroll
        for (int i = 0; i < SR; ++i)
#pragma unroll
            for (int j = 0; j < SC; ++j)
                qv[i][j] = (valid >> (i * SC + j)) & 1u
                               ? as_global(bd.q)[lane_off + (long long)16 * i * img.W + 64 * j]
                               : 0.0f;

        const int dl = tid & 63;
        const int tx = (dl >> 5) * LH + (dl & (LH - 1)), ty = (tid >> 6) * P + ((dl & 31) / LH);
        const int px0 = x0 + tx * P, py = y0 + ty;
        const float *lane_m = mt + (ty + R) * lp + tx * P + R;
        float dx[P];
#pragma unroll
        for (int p = 0; p < P; ++p) dx[p] = 0.0f;

        for (int d = 0; d < D; ++d) {
            // Recomputed per direction rather than held: the slot masks and
            // addresses derived from these would otherwise stay live across
            // the walk.
            unsigned vb = valid;
            long long off = lane_off;
            asm volatile("" : "+v"(vb), "+v"(off));
            const auto mp = as_global(img.map) + (size_t)d * plane;
            int sv[SR][SC];
#pragma unroll
            for (int i = 0; i < SR; ++i)
#pragma unroll
                for (int j = 0; j < SC; ++j)
                    sv[i][j] = (vb >> (i * SC + j)) & 1u ? (int)mp[off + (long long)16 * i * img.W + 64 * j] : -1;
            if constexpr (RING) {
                // The deepest scale selected in the window (elements outside
                // it hold -1): the walk starts there.
                int deepest = -1;
#pragma unroll
                for (int i = 0; i < SR; ++i)
#pragma unroll
                    for (int j = 0; j < SC; ++j) deepest = max(deepest, sv[i][j]);
                const int top = block_max<DXT>(deepest);  // its first barrier: the last ring sums are done
                if (top < 0) continue;
                // Clear the map, then each step writes only the elements that
                // join it; elements outside the window hold scale -1 and never
                // join.
                for (int e = tid; e < map4; e += DXT) mt4[e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                for (int s = top; s >= 0; --s) {
                    __syncthreads();  // the clear, or the previous ring sums, done with the map
                    const int k = d * S + s;
                    const float cw = as_const(inv_t)[k] * as_const(rings.scale)[k];
#pragma unroll
                    for (int i = 0; i < SR; ++i)
#pragma unroll
                        for (int j = 0; j < SC; ++j)
                            if (sv[i][j] == s) mt[(r0 + 16 * i) * lp + c0 + 64 * j] = qv[i][j] * cw;
                    __syncthreads();
                    ring_sums_sliding<P>(lane_m, lp, rings.flip, load_const(rings.kern + k), dx);
                }
                // the next direction's block_max is the barrier before its clear
            } else {
                for (int s = 0; s < S; ++s) {
                    bool any = false;
                    __syncthreads();  // previous correlation done with the map
                    const float inv = as_const(inv_t)[d * S + s];
#pragma unroll
                    for (int i = 0; i < SR; ++i)
#pragma unroll
                        for (int j = 0; j < SC; ++j) {
                            const int r = r0 + 16 * i, c = c0 + 64 * j;
                            if (r < fh && c < fw) {
                                const bool m = sv[i][j] == s && qv[i][j] != 0.0f;
                                any |= m;
                                mt[r * lp + c] = m ? qv[i][j] * inv : 0.0f;
                            }
                        }
                    if (!block_vote<DXT, false>(any)) continue;
                    correlate<P, false, 0>(lane_m, lp, adj.flip, d * S + s, dx);
                }
            }
        }
        const auto out = as_global(bd.dx) + (size_t)py * img.pitch + px0;
#pragma unroll
        for (int p = 0; p < P; ++p)
            if (py < img.H && px0 + p < img.W) out[p] = dx[p];
        __syncthreads();  // LDS and `next` are reused by the next tile
    }
}

// Pass 3a: column sums of the partial rows, in double:
// sums[c] = sum_b partial[b][c] for the NW tap columns (A) and the K kernel
// columns (E).  dL/dw = A - 2 w E / sumsq is a difference of two terms that
// nearly cancel for some taps, so both reach the final subtraction with no
// more rounding than the per-row float sums hold.  64 columns per work-group
// (consecutive lanes read consecutive columns), 16 row groups summed in a
// fixed order.
constexpr int kSumCols = 64, kSumRowGroups = 16;
__global__ __launch_bounds__(kSumCols * kSumRowGroups) void column_sums_kernel(const float *__restrict__ partial,
                                                                               int rows, int cols,
                                                                               double *__restrict__ sums)
{
    __shared__ double part[kSumRowGroups][kSumCols];
    const int tx = threadIdx.x % kSumCols, ty = threadIdx.x / kSumCols;
    const int c = blockIdx.x * kSumCols + tx;
    double s = 0.0;
    if (c < cols)
        for (int b = ty; b < rows; b += kSumRowGroups) s += (double)partial[(size_t)b * cols + c];
    part[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && c < cols) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kSumRowGroups; ++i) t += part[i][tx];
        sums[c] = t;
    }
}

// Pass 3b: dkernel[i] += A[i] - 2 w[i] E_k / sumsq_k, in double until the one
// rounding to float at the end.
__global__ __launch_bounds__(kThreads) void combine_kernel(const double *__restrict__ sums, int NW,
                                                           const float *__restrict__ weights,
                                                           const int *__restrict__ tap_kernel,
                                                           const float *__restrict__ sumsq,
                                                           float *__restrict__ dkernel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NW) return;
    const int k = tap_kernel[i];
    dkernel[i] = (float)((double)dkernel[i] + (sums[i] - 2.0 * (double)weights[i] * sums[NW + k] / (double)sumsq[k]));
}

size_t grad_lds_bytes(int R, int K) { return sizeof(float) * ((size_t)grad_tab_offset(R) + 3 * K); }
size_t dx_lds_bytes(int R, bool ring)
{
    // (H + 2R) rows of pitch DXLP, in whole float4s: the ring form clears the
    // map 16 bytes at a time
    const size_t floats = (size_t)((ring ? DxShape<true>::H : DxShape<false>::H) + 2 * R) * DXLP;
    return sizeof(float4) * ((floats + 3) / 4);
}

// The dx pass's LDS (the 64 x 128 map of the ring form) can pass the default
// limit of a launch: raise it.
hipError_t allow_dx_lds(int R)
{
    for (bool ring : {true, false}) {
        const size_t bytes = dx_lds_bytes(R, ring);
        if (bytes <= 65536) continue;
        const void *f = ring ? reinterpret_cast<const void *>(&bwd_dx_kernel<true>)
                             : reinterpret_cast<const void *>(&bwd_dx_kernel<false>);
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The first pass's LDS can pass the default limit of a launch: raise it.
hipError_t allow_grad_lds(int R, int K)
{
    const size_t bytes = grad_lds_bytes(R, K);
    if (bytes <= 65536) return hipSuccess;
    for (const void *f : {reinterpret_cast<const void *>(&bwd_grad_kernel<true>),
                          reinterpret_cast<const void *>(&bwd_grad_kernel<false>)}) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

void backward_occupancy(int radius, bool ring, int num_kernels, int *prep, int *dx)
{
    int b = 0;
    if (allow_grad_lds(radius, num_kernels) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bwd_grad_kernel<true>, GNT,
                                                     grad_lds_bytes(radius, num_kernels)) != hipSuccess)
        b = 1;
    *prep = b > 0 ? b : 1;
    b = 0;
    const void *f = ring ? reinterpret_cast<const void *>(&bwd_dx_kernel<true>)
                         : reinterpret_cast<const void *>(&bwd_dx_kernel<false>);
    if (allow_dx_lds(radius) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, DXT, dx_lds_bytes(radius, ring)) != hipSuccess)
        b = 1;
    *dx = b > 0 ? b : 1;
}

hipError_t launch_backward(const DevKset &ks, const DevRings &rings, const DevAdjoint &adj, const BwdDesc *d_imgs,
                           const uint2 *d_tiles,
                           int num_tiles, const uint2 *dx_tiles, int num_dx_tiles, int *d_counter, int radius,
                           int grid_prep, int grid_dx, float *partial, float *dkernel, hipStream_t stream)
{
    const int K = ks.D * ks.S, NW = adj.num_weights;
    if (ks.S > kMaxScales) return hipErrorInvalidValue;
    const bool ring = rings.runs != nullptr;
    if (grad_lds_bytes(radius, K) > (size_t)kMaxLdsBytes || dx_lds_bytes(radius, ring) > (size_t)kMaxLdsBytes)
        return hipErrorInvalidValue;
    hipError_t ea = allow_grad_lds(radius, K);
    if (ea == hipSuccess) ea = allow_dx_lds(radius);
    if (ea != hipSuccess) return ea;
    const int rows = grid_prep;
    const size_t row_floats = (size_t)rows * (NW + K);
    hipLaunchKernelGGL(reset_counter2, dim3(1), dim3(1), 0, stream, d_counter);
    hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(kThreads), 0, stream, partial, row_floats, 0.0f);
    if (rings.runs)
        hipLaunchKernelGGL(bwd_grad_kernel<true>, dim3(grid_prep), dim3(GNT), grad_lds_bytes(radius, K), stream,
                           ks, rings, adj, d_imgs, d_tiles, num_tiles, d_counter, radius, partial);
    else
        hipLaunchKernelGGL(bwd_grad_kernel<false>, dim3(grid_prep), dim3(GNT), grad_lds_bytes(radius, K), stream,
                           ks, rings, adj, d_imgs, d_tiles, num_tiles, d_counter, radius, partial);
    if (ring)
        hipLaunchKernelGGL(bwd_dx_kernel<true>, dim3(grid_dx), dim3(DXT), dx_lds_bytes(radius, true), stream, ks,
                           rings, adj, d_imgs, dx_tiles, num_dx_tiles, d_counter, radius);
    else
        hipLaunchKernelGGL(bwd_dx_kernel<false>, dim3(grid_dx), dim3(DXT), dx_lds_bytes(radius, false), stream, ks,
                           rings, adj, d_imgs, dx_tiles, num_dx_tiles, d_counter, radius);
    if (dkernel) {
        // NW + K doubles after the rows, at the next 8-byte boundary
        double *sums = (double *)(partial + ((row_floats + 1) & ~(size_t)1));
        const int cols = NW + K;
        hipLaunchKernelGGL(column_sums_kernel, dim3((cols + kSumCols - 1) / kSumCols), dim3(kSumCols * kSumRowGroups),
                           0, stream, partial, rows, cols, sums);
        hipLaunchKernelGGL(combine_kernel, dim3((NW + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, sums, NW,
                           ks.weights, adj.tap_kernel, adj.sumsq, dkernel);
    }
    return hipGetLastError();
}

}  // namespace lpa
