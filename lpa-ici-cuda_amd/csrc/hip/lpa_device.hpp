// Device-side data layout shared by the forward and adjoint kernels.
//
// Kernel set in HBM (run-length, row runs), built once per set:
//   runs[r]    = {row offset from the kernel centre, first column offset,
//                 run length, index of the run's first weight}
//   weights[]  = the runs' weights, run after run (interior zeros kept)
//   kern[k]    = {first run, number of runs} for kernel k = d*S + s
//   runs_b[r]  = run r for an LDS tile of one row pitch lp, in bytes:
//                {4 (row lp + col), run length, 4 x first weight index, 0}
//                (one table per pitch in use; null where none was built)
// Runs are in the dense kernel's row-major order (the reference's tap order).
// Per-image tables, [3][K] floats: variance, ICI threshold and 1/variance,
// each computed on the host in float exactly as the reference does.
#ifndef LPA_DEVICE_HPP
#define LPA_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpa {

struct DevKset {
    const int4 *runs;
    const float *weights;
    const int2 *kern;
    const int4 *runs_b;
    int D, S;
};

// Ring form of a kernel set whose kernels have uniform weights and nested
// supports across scales (every scale's support contains the previous
// one's): kernel (d,s) = scale[k] * (indicator of its support), and the
// support is the union of the rings (support(d,s) minus support(d,s-1)) up to
// s.  runs/kern describe the rings as row runs {row, col, len, 0}; a
// scale's estimate is scale[k] times the running sum over its rings.
struct DevRings {
    const int4 *runs;
    const int4 *flip;  // the ring runs rotated by 180 degrees (same kern ranges)
    const int2 *kern;
    // The ring runs for a float tile of one row pitch, in bytes: {4 (row lp +
    // col), run length, 0, 0} (null where none was built).
    const int4 *runs_b;
    // The ring runs for a fixed-point prefix table of one row pitch lp, as
    // byte offsets from a lane's first prefix value: pairs[i] = {a0, b0, a1,
    // b1} for two consecutive runs of a kernel, a = 4 (row lp + col) and
    // b = a + 4 len (a run adds Q[b] - Q[a]); pair_kern[k] = {first pair,
    // number of runs} (an odd kernel's last pair uses only a0, b0).
    const int4 *pairs;
    const int2 *pair_kern;
    const float *scale;    // w_s per kernel
    const float *scale_q;  // w_s * 2^-F per kernel (exact: a power-of-two scaling)
    float qscale;  // 2^F: fixed-point scale of tile values for ring sums
    float qrange;  // windows whose values all lie strictly within +-qrange use fixed point
};

// Arithmetic of the forward walk: FMA over the weighted runs, the
// reference's separately rounded mul/add, ring sums in fixed point (float
// ring sums for a window outside the fixed-point range), float ring sums.
enum WalkMode { kWalkFma = 0, kWalkExact = 1, kWalkRing = 2, kWalkRingFloat = 3 };

// One image of a batch.  tab points at its [3][K] table.
struct ImgDesc {
    const float *x;
    float *y;
    uint8_t *map;
    const float *tab;
    int W, H, pitch;
    int tiles_x;
    float *est;  // [D][H][W] selected estimates (forward output, adjoint input), may be null
};

// Compiled tile shapes: output tile tw x th pixels, px adjacent pixels per
// lane along a row, nt lanes per work-group, the inner-loop body (2: the
// chain of unrolled taps with exits, 0: sliding-window loop), the LDS row
// pitch's residue mod 32 (the bank offset between consecutive rows), the
// lane layout (split 1: each half-wave, whose 32 lanes the LDS serves
// together, takes its own half of the rows' lanes, so the 32 lanes of a read
// sit 8 or 4 px apart on rows one bank apart: 32 distinct banks) and the
// waves per SIMD the register budget is held to (wpe; 1: no limit).
struct TileShape {
    int tw, th, px, body, pm, nt, split, wpe;
};
constexpr int kThreads = 256;
constexpr int kNumShapes = 10;
constexpr TileShape kShapes[kNumShapes] = {
    {32, 32, 4, 2, 1, 256, 0, 1},   // 0: wave covers 32x8 px
    {32, 32, 4, 0, 1, 256, 0, 1},   // 1: same, window loop
    {64, 16, 4, 2, 1, 256, 0, 1},   // 2: wave covers 64x4 px
    {32, 16, 2, 2, 1, 256, 0, 1},   // 3: wave covers 32x4 px
    {64, 32, 8, 2, 1, 256, 0, 1},   // 4: wave covers 64x8 px
    {32, 64, 8, 2, 1, 256, 0, 1},   // 5: wave covers 32x16 px; 4 lanes per row keep 8-px lanes in distinct banks
    {32, 64, 8, 2, 2, 256, 0, 1},   // 6: as 5, rows two banks apart: the dword pairs of 16 lanes (4 rows) are disjoint
    {64, 64, 8, 2, 1, 512, 1, 4},   // 7: halo 3.9 words per output pixel against 5.8; two work-groups per CU
    {64, 32, 4, 2, 1, 512, 1, 6},   // 8: three work-groups of 512 lanes per CU (24 waves)
    {64, 64, 4, 2, 1, 1024, 1, 8},  // 9: 1024 lanes within 64 VGPRs: two work-groups per CU (32 waves)
};

constexpr int kMaxRadius = 31;
// LDS of one work-group on gfx950.
constexpr int kMaxLdsBytes = 160 * 1024;

// LDS row pitch for a tile of width tw and halo radius r: at least tw + 2r and
// congruent to pm mod 32 (default 1: lanes four pixels apart on consecutive
// rows fall in different banks).
__host__ __device__ constexpr int lds_pitch(int tw, int r, int pm = 1)
{
    const int p = tw + 2 * r;
    return p + ((pm - p) % 32 + 32) % 32;
}
__host__ __device__ constexpr int lds_floats(int tw, int th, int r, int pm = 1)
{
    return (th + 2 * r) * lds_pitch(tw, r, pm);
}
// Row pitch of the fixed-point prefix table of a tw-wide tile (one leading
// column, Q[r][0] = 0).
__host__ __device__ constexpr int prefix_pitch(int tw, int r, int pm = 1) { return lds_pitch(tw + 1, r, pm); }

// One image of an adjoint batch: the forward's image (x, y = the forward's
// output, map = its scale indices, tab) plus the upstream gradient g, the
// output dx and a per-pixel workspace q = g / sum_d(1/var_d).
struct BwdDesc {
    ImgDesc img;
    const float *g;
    float *dx;
    float *q;
};

// Run pieces the adjoint's tap walk is unrolled for (longer runs are walked
// in pieces of this length).
constexpr int kGradPiece = 8;

// Extra kernel-set tables the adjoint reads.
//   flip: the 180-degree rotated runs (-row, -(col + len - 1), len) with
//         reversed weights, so the transposed correlation reuses correlate()
//   tap_lds[i] = row * lp + col for run weight i at offset (row, col) from the
//                kernel centre: its offset in the adjoint's staged tile, whose
//                row pitch is lp = lds_pitch(kBwdTileW, radius, kBwdPitchResidue)
//   tap_kernel[i] = kernel index k of run weight i
//   kern_w[k] = {first run weight, number of run weights, pieces, 0} of
//               kernel k; pieces = sum over its runs of ceil(len / kGradPiece)
struct DevAdjoint {
    DevKset flip;
    const int *tap_lds;
    const int *tap_kernel;
    const int4 *kern_w;
    const float *sumsq;
    int num_weights;
};

// Launch entry points (lpa_fwd.hip).
hipError_t launch_forward_plain(int shape, WalkMode mode, const DevKset &ks, const DevRings &rings,
                                const ImgDesc &img, int radius, hipStream_t stream);
hipError_t launch_forward_batched(int shape, WalkMode mode, const DevKset &ks, const DevRings &rings,
                                  const ImgDesc *d_imgs, const uint2 *d_tiles, int num_tiles, int *d_counter,
                                  int radius, int grid, hipStream_t stream);
// Work-groups of one shape that fit on a CU at this radius and kernel count.
int blocks_per_cu(int shape, int radius, int num_kernels);

// Per-direction estimates at frozen scales (lpa_fwd.hip): est = [D][H][W]
// floats per image, dense like the scale map; a persistent grid over the tile
// queue.  The fixed-point ring walk runs on the forward's tile configuration
// (`shape`: its choice of fixed point or float per tile window must be the
// forward's); every other walk gives the same bits on any configuration and
// runs on kEstimateShape.
struct EstDesc {
    ImgDesc img;
    float *est;
};
constexpr int kEstimateShape = 5;
hipError_t launch_estimates(WalkMode mode, int shape, const DevKset &ks, const DevRings &rings,
                            const EstDesc *d_imgs, const uint2 *d_tiles, int num_tiles, int *d_counter, int radius,
                            int grid, hipStream_t stream);

// Adjoint (lpa_bwd.hip).  The first pass walks 64 x 64 tiles (d_tiles), the
// dx pass tiles 64 wide and kDxTileRowsRing (ring form) or
// kDxTileRowsWeighted tall (dx_tiles); both queues are laid out as for the
// batched forward ({image, (tile_y << 16) | tile_x}).
//   partial: [grid_prep][num_weights + K] floats (+1+2(num_weights+K) for the column sums in double), one row
//            per work-group of the first pass, accumulated by its waves
//   dkernel: += reduction of the rows plus the variance term (may be null)
constexpr int kBwdTileW = 64, kBwdTileH = 64, kDxTileW = 64, kDxTileRowsRing = 128, kDxTileRowsWeighted = 64;
// Row pitch residue (mod 32 banks) of the first pass's float tile: with rows
// 9 banks apart, the 64 taps a wave reads together when the taps of a kernel
// are on the lanes spread over the banks for every direction of the shipped
// sets about twice as well as with rows one bank apart.
constexpr int kBwdPitchResidue = 9;
//   rings:   the ring form (runs == nullptr when the set has none); the
//            forward's default walk, which the adjoint repeats for yhat
hipError_t launch_backward(const DevKset &ks, const DevRings &rings, const DevAdjoint &adj, const BwdDesc *d_imgs,
                           const uint2 *d_tiles,
                           int num_tiles, const uint2 *dx_tiles, int num_dx_tiles, int *d_counter, int radius,
                           int grid_prep, int grid_dx, float *partial, float *dkernel, hipStream_t stream);
// Work-groups per CU of the two tile passes.
void backward_occupancy(int radius, bool ring, int num_kernels, int *prep, int *dx);

}  // namespace lpa

#endif
