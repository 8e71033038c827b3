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
// Pass 1 (prep, one tile at a time): q, the recomputed yhat of the selected
// scale, E and A.  Per direction the tile's pixels are grouped by selected
// scale (deterministic ballot order); A is formed with one lane per tap,
// reducing over the group, into the work-group's own row of `partial`.
// Pass 2 (dx): per direction and scale, the masked map q*inv*[s_d = s] over
// tile + halo is correlated with the 180-degree rotated kernel.
// Pass 3: rows of `partial` are summed and the variance term is added.
#include <hip/hip_runtime.h>

#include "lpa_device.hpp"
#include "lpa_tile.hpp"

namespace lpa {
namespace {

constexpr int TW = 32, TH = 32, P = 4, LX = TW / P;
constexpr int kTilePx = TW * TH;
constexpr int kMaxScales = 16;

struct PrepLds {
    // Offsets (bytes) of each region inside the dynamic LDS block.
    int x, tab, q, e, sel, list_q, list_b, cnt, eblk, total;
};

__host__ __device__ inline PrepLds prep_layout(int R, int S, int K)
{
    PrepLds l;
    int o = 0;
    l.x = o;
    o += 4 * lds_floats(TW, TH, R);
    l.tab = o;
    o += 4 * 3 * K;
    l.q = o;
    o += 4 * kTilePx;
    l.e = o;
    o += 4 * kTilePx;
    l.list_q = o;
    o += 4 * kTilePx;
    l.list_b = o;
    o += 4 * kTilePx;
    l.sel = o;
    o += kTilePx;
    l.cnt = o;
    o += 4 * 2 * kMaxScales;
    l.eblk = o;
    o += 4 * K;
    l.total = (o + 15) & ~15;
    (void)S;
    return l;
}

__device__ __forceinline__ float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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

// Pass 1.  counter[0] feeds the tile queue.
__global__ __launch_bounds__(kThreads) void bwd_prep_kernel(DevKset ks, DevAdjoint adj,
                                                            const BwdDesc *__restrict__ imgs,
                                                            const uint2 *__restrict__ tiles, int num_tiles,
                                                            int *__restrict__ counter, int R,
                                                            float *__restrict__ partial)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next;
    const int S = ks.S, D = ks.D, K = D * S;
    const PrepLds L = prep_layout(R, S, K);
    float *xt = (float *)(lds + L.x);
    float *tab = (float *)(lds + L.tab);
    float *qt = (float *)(lds + L.q);
    float *et = (float *)(lds + L.e);
    float *lq = (float *)(lds + L.list_q);
    int *lb = (int *)(lds + L.list_b);
    uint8_t *sel = lds + L.sel;
    int *cnt = (int *)(lds + L.cnt);  // [0, S): counts, [kMaxScales, +S): offsets
    float *eblk = (float *)(lds + L.eblk);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lp = lds_pitch(TW, R);
    const int NW = adj.num_weights;
    float *__restrict__ row = partial + (size_t)blockIdx.x * (NW + K);

    for (int k = tid; k < K; k += kThreads) eblk[k] = 0.0f;
    for (;;) {
        if (tid == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 te = tiles[t];
        const BwdDesc bd = imgs[te.x];
        const ImgDesc &img = bd.img;
        const int x0 = (int)(te.y & 0xffffu) * TW, y0 = (int)(te.y >> 16) * TH;
        for (int i = tid; i < 3 * K; i += kThreads) tab[i] = img.tab[i];
        stage_tile<TW, TH>(xt, lp, img.x, img.W, img.H, img.pitch, x0, y0, R);
        __syncthreads();

        const float *inv_t = tab + 2 * K;
        const int tx = tid % LX, ty = tid / LX;
        const int px0 = x0 + tx * P, py = y0 + ty;
        const float *lane_x = xt + (ty + R) * lp + tx * P + R;
        const size_t plane = (size_t)img.W * img.H;
        const size_t pix = (size_t)py * img.W + px0;
        bool inside[P];
        float q[P], yv[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            inside[p] = py < img.H && px0 + p < img.W;
            float wt = 0.0f;
            for (int d = 0; d < D; ++d) {
                const int s = inside[p] ? min((int)img.map[(size_t)d * plane + pix + p], S - 1) : 0;
                wt += inv_t[d * S + s];
            }
            const size_t gi = (size_t)py * img.pitch + px0 + p;
            q[p] = inside[p] ? bd.g[gi] / wt : 0.0f;
            yv[p] = inside[p] ? img.y[gi] : 0.0f;
            qt[ty * TW + tx * P + p] = q[p];
            if (inside[p]) bd.q[(size_t)py * img.W + px0 + p] = q[p];
        }

        for (int d = 0; d < D; ++d) {
            int sd[P];
#pragma unroll
            for (int p = 0; p < P; ++p) sd[p] = inside[p] ? img.map[(size_t)d * plane + pix + p] : 255;
            // Recompute yhat_d at the selected scale; a wave runs a scale
            // only if one of its pixels selected it.
            float yh[P];
#pragma unroll
            for (int p = 0; p < P; ++p) yh[p] = 0.0f;
            for (int s = 0; s < S; ++s) {
                bool want = false;
#pragma unroll
                for (int p = 0; p < P; ++p) want |= sd[p] == s;
                if (!__any(want)) continue;
                float acc[P];
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = 0.0f;
                if (want) correlate<P, false>(lane_x, lp, ks, d * S + s, acc);
#pragma unroll
                for (int p = 0; p < P; ++p)
                    if (sd[p] == s) yh[p] = acc[p];
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const int i = ty * TW + tx * P + p;
                const int si = sd[p] < S ? sd[p] : 0;
                et[i] = inside[p] ? q[p] * inv_t[d * S + si] * (yh[p] - yv[p]) : 0.0f;
                sel[i] = (uint8_t)sd[p];
            }
            __syncthreads();

            // Group the tile's pixels by selected scale: counts, offsets,
            // then placement in ballot order.  Wave w owns scales w, w+4, ...
            for (int s = wave; s < S; s += kThreads / 64) {
                int c = 0;
                float esum = 0.0f;
                for (int ch = 0; ch < kTilePx / 64; ++ch) {
                    const int i = ch * 64 + lane;
                    const bool m = sel[i] == s;
                    c += __popcll(__ballot(m));
                    esum += m ? et[i] : 0.0f;
                }
                esum = wave_sum(esum);
                if (lane == 0) {
                    cnt[s] = c;
                    eblk[d * S + s] += esum;
                }
            }
            __syncthreads();
            if (tid == 0) {
                int o = 0;
                for (int s = 0; s < S; ++s) {
                    cnt[kMaxScales + s] = o;
                    o += cnt[s];
                }
            }
            __syncthreads();
            for (int s = wave; s < S; s += kThreads / 64) {
                int pos = cnt[kMaxScales + s];
                const float inv = inv_t[d * S + s];
                for (int ch = 0; ch < kTilePx / 64; ++ch) {
                    const int i = ch * 64 + lane;
                    const bool m = sel[i] == s;
                    const unsigned long long mask = __ballot(m);
                    if (m) {
                        const int at = pos + __popcll(mask & ((1ull << lane) - 1ull));
                        lq[at] = qt[i] * inv;
                        lb[at] = ((i / TW) + R) * lp + (i % TW) + R;
                    }
                    pos += __popcll(mask);
                }
            }
            __syncthreads();

            // A_k[t] += sum over the group of c_d(p) x[p+t], one lane per tap.
            for (int s = 0; s < S; ++s) {
                const int n = cnt[s];
                if (n == 0) continue;
                const int k = d * S + s;
                const int2 kw = adj.kern_w[k];
                const float *gq = lq + cnt[kMaxScales + s];
                const int *gb = lb + cnt[kMaxScales + s];
                for (int j = tid; j < kw.y; j += kThreads) {
                    const int2 tp = adj.tap_pos[kw.x + j];
                    const int off = tp.x * lp + tp.y;
                    float a0 = 0.0f, a1 = 0.0f;
                    int e = 0;
                    for (; e + 2 <= n; e += 2) {
                        a0 = __builtin_fmaf(gq[e], xt[gb[e] + off], a0);
                        a1 = __builtin_fmaf(gq[e + 1], xt[gb[e + 1] + off], a1);
                    }
                    if (e < n) a0 = __builtin_fmaf(gq[e], xt[gb[e] + off], a0);
                    row[kw.x + j] += a0 + a1;
                }
            }
        }
        __syncthreads();  // LDS regions and `next` are reused by the next tile
    }
    __syncthreads();
    for (int k = tid; k < K; k += kThreads) row[NW + k] = eblk[k];
}

// Pass 2 (dx), on 64 x 64 tiles with 1024 lanes (4 px per lane).  Each lane
// keeps q and the scale index of 16 tile+halo elements in registers (rows
// tid/64 + 16i, columns tid%64 + 64j); the LDS holds only the masked map and
// the table.  The wide tile cuts the masked-map building per output pixel
// (halo overhead (64+2R)^2/64^2 instead of (32+2R)^2/32^2).
constexpr int DXW = 64, DXH = 64, DXP = 4, DXT = 1024, DXLX = DXW / DXP;
constexpr int kSlotRows = 8, kSlotCols = 2;  // 16 + 16i rows, 64 + 64j columns cover 126 x 126

__global__ __launch_bounds__(DXT) void bwd_dx_kernel(DevKset ks, DevAdjoint adj, const BwdDesc *__restrict__ imgs,
                                                     const uint2 *__restrict__ tiles, int num_tiles,
                                                     int *__restrict__ counter, int R)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next;
    const int S = ks.S, D = ks.D, K = D * S;
    const int lp = lds_pitch(DXW, R);
    float *mt = (float *)lds;
    const int tid = threadIdx.x;
    const int fw = DXW + 2 * R, fh = DXH + 2 * R;
    const int r0 = tid >> 6, c0 = tid & 63;

    for (;;) {
        if (tid == 0) next = atomicAdd(counter + 1, 1);
        __syncthreads();
        const int t = next;
        if (t >= num_tiles) break;
        const uint2 te = tiles[t];
        const BwdDesc bd = imgs[te.x];
        const ImgDesc &img = bd.img;
        const int x0 = (int)(te.y & 0xffffu) * DXW, y0 = (int)(te.y >> 16) * DXH;
        const float *__restrict__ inv_t = img.tab + 2 * K;  // wave-uniform scalar loads
        const size_t plane = (size_t)img.W * img.H;

        float qv[kSlotRows][kSlotCols];
#pragma unroll
        for (int i = 0; i < kSlotRows; ++i)
#pragma unroll
            for (int j = 0; j < kSlotCols; ++j) {
                const int r = r0 + 16 * i, c = c0 + 64 * j;
                const int gy = y0 - R + r, gx = x0 - R + c;
                const bool in = r < fh && c < fw && gy >= 0 && gy < img.H && gx >= 0 && gx < img.W;
                qv[i][j] = in ? bd.q[(size_t)gy * img.W + gx] : 0.0f;
            }

        const int tx = tid % DXLX, ty = tid / DXLX;
        const int px0 = x0 + tx * DXP, py = y0 + ty;
        const float *lane_m = mt + (ty + R) * lp + tx * DXP + R;
        float dx[DXP];
#pragma unroll
        for (int p = 0; p < DXP; ++p) dx[p] = 0.0f;

        for (int d = 0; d < D; ++d) {
            int sv[kSlotRows][kSlotCols];
#pragma unroll
            for (int i = 0; i < kSlotRows; ++i)
#pragma unroll
                for (int j = 0; j < kSlotCols; ++j) {
                    const int r = r0 + 16 * i, c = c0 + 64 * j;
                    const int gy = y0 - R + r, gx = x0 - R + c;
                    const bool in = r < fh && c < fw && gy >= 0 && gy < img.H && gx >= 0 && gx < img.W;
                    sv[i][j] = in ? img.map[(size_t)d * plane + (size_t)gy * img.W + gx] : 255;
                }
            for (int s = 0; s < S; ++s) {
                int any = 0;
                __syncthreads();  // previous correlation done with the map
                const float inv = inv_t[d * S + s];
#pragma unroll
                for (int i = 0; i < kSlotRows; ++i)
#pragma unroll
                    for (int j = 0; j < kSlotCols; ++j) {
                        const int r = r0 + 16 * i, c = c0 + 64 * j;
                        if (r < fh && c < fw) {
                            const bool m = sv[i][j] == s && qv[i][j] != 0.0f;
                            any |= m;
                            mt[r * lp + c] = m ? qv[i][j] * inv : 0.0f;
                        }
                    }
                if (!__syncthreads_or(any)) continue;
                correlate<DXP, false, 0>(lane_m, lp, adj.flip, d * S + s, dx);
            }
        }
        float *__restrict__ out = bd.dx + (size_t)py * img.pitch + px0;
#pragma unroll
        for (int p = 0; p < DXP; ++p)
            if (py < img.H && px0 + p < img.W) out[p] = dx[p];
        __syncthreads();  // LDS and `next` are reused by the next tile
    }
}

// Pass 3a: E_k = sum of the rows' E entries.  One thread per kernel.
__global__ __launch_bounds__(kThreads) void reduce_e_kernel(const float *__restrict__ partial, int rows, int NW,
                                                            int K, float *__restrict__ e_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    float s = 0.0f;
    for (int b = 0; b < rows; ++b) s += partial[(size_t)b * (NW + K) + NW + k];
    e_out[k] = s;
}

// Pass 3b: dkernel[i] += sum_b A_b[i] - 2 w[i] E_k / sumsq_k.
__global__ __launch_bounds__(kThreads) void reduce_a_kernel(const float *__restrict__ partial, int rows, int NW,
                                                            int K, const float *__restrict__ e_k,
                                                            const float *__restrict__ weights,
                                                            const int *__restrict__ tap_kernel,
                                                            const float *__restrict__ sumsq,
                                                            float *__restrict__ dkernel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NW) return;
    float s = 0.0f;
    for (int b = 0; b < rows; ++b) s += partial[(size_t)b * (NW + K) + i];
    const int k = tap_kernel[i];
    dkernel[i] += s - 2.0f * weights[i] * e_k[k] / sumsq[k];
}

size_t prep_lds_bytes(int R, int S, int K) { return (size_t)prep_layout(R, S, K).total; }
size_t dx_lds_bytes(int R, int K)
{
    (void)K;
    return sizeof(float) * (size_t)lds_floats(DXW, DXH, R);
}

}  // namespace

void backward_occupancy(int radius, int num_scales, int num_kernels, int *prep, int *dx)
{
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bwd_prep_kernel, kThreads,
                                                     prep_lds_bytes(radius, num_scales, num_kernels)) != hipSuccess)
        b = 1;
    *prep = b > 0 ? b : 1;
    b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bwd_dx_kernel, DXT, dx_lds_bytes(radius, num_kernels)) !=
        hipSuccess)
        b = 1;
    *dx = b > 0 ? b : 1;
}

hipError_t launch_backward(const DevKset &ks, const DevAdjoint &adj, const BwdDesc *d_imgs, const uint2 *d_tiles,
                           int num_tiles, const uint2 *dx_tiles, int num_dx_tiles, int *d_counter, int radius,
                           int grid_prep, int grid_dx, float *partial, float *dkernel, hipStream_t stream)
{
    const int K = ks.D * ks.S, NW = adj.num_weights;
    if (ks.S > kMaxScales) return hipErrorInvalidValue;
    if (prep_lds_bytes(radius, ks.S, K) > 65536 || dx_lds_bytes(radius, K) > 65536) return hipErrorInvalidValue;
    const size_t rows = (size_t)grid_prep * (NW + K);
    hipLaunchKernelGGL(reset_counter2, dim3(1), dim3(1), 0, stream, d_counter);
    hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(kThreads), 0, stream, partial, rows, 0.0f);
    hipLaunchKernelGGL(bwd_prep_kernel, dim3(grid_prep), dim3(kThreads), prep_lds_bytes(radius, ks.S, K), stream, ks,
                       adj, d_imgs, d_tiles, num_tiles, d_counter, radius, partial);
    hipLaunchKernelGGL(bwd_dx_kernel, dim3(grid_dx), dim3(DXT), dx_lds_bytes(radius, K), stream, ks, adj, d_imgs,
                       dx_tiles, num_dx_tiles, d_counter, radius);
    if (dkernel) {
        float *e_k = partial + rows;  // K floats after the rows
        hipLaunchKernelGGL(reduce_e_kernel, dim3((K + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, partial,
                           grid_prep, NW, K, e_k);
        hipLaunchKernelGGL(reduce_a_kernel, dim3((NW + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, partial,
                           grid_prep, NW, K, e_k, ks.weights, adj.tap_kernel, adj.sumsq, dkernel);
    }
    return hipGetLastError();
}

}  // namespace lpa
