// C ABI over the HIP kernels: kernel sets, plans and the forward pass.
// See include/lpa_ici.h for the contract of every entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../internal.hpp"
#include "lpa_device.hpp"
#include "lpa_ici.h"
#include "lpa_tile.hpp"

using lpa::fail;

// Workspace of lpa_ici_forward_noalloc: the work counter, image descriptors,
// per-image tables and tile queue of one call, in a fixed array that every
// device holds for the life of the process.  Host code must never pass the
// array's name to a kernel or a copy (on the host it is only a stand-in for
// the symbol): it uses the device address hipGetSymbolAddress returns, looked
// up once per device (scratch_address).
constexpr size_t kStaticWorkspaceBytes = 4u << 20;
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

struct lpa_kset {
    int D = 0, S = 0, K = 0;
    int radius = 0;
    int num_nonzero = 0;
    std::vector<int4> runs;
    std::vector<float> weights;
    std::vector<int2> kern;
    std::vector<int> nnz;
    std::vector<float> sumsq;
    int device = 0;
    int4 *d_runs = nullptr;
    float *d_weights = nullptr;
    int2 *d_kern = nullptr;
    // Stretches of consecutive runs of one length (see lpa::DevKset).
    int shape = -1;  // -1: built-in choice
    // Adjoint tables (see lpa::DevAdjoint).
    std::vector<int4> runs_flip;
    std::vector<float> weights_flip;
    std::vector<int> tap_lds;
    std::vector<int4> kern_w;
    std::vector<int> tap_kernel;
    int4 *d_runs_flip = nullptr;
    float *d_weights_flip = nullptr;
    int *d_tap_lds = nullptr;
    int4 *d_kern_w = nullptr;
    int *d_tap_kernel = nullptr;
    float *d_sumsq = nullptr;
    // Ring form (see lpa::DevRings), when every kernel has uniform weights
    // and nested supports across scales.
    bool has_rings = false;
    std::vector<int4> ring_runs;
    std::vector<int2> ring_kern;
    std::vector<float> ring_scale;
    std::vector<int4> ring_runs_flip;  // the rings rotated by 180 degrees (adjoint)
    int ring_taps = 0;
    int ring_frac_bits = 0;      // F: ring sums take tile values as rint(v * 2^F)
    float ring_range = 0.0f;     // fixed point for windows with every |v| below this
    int4 *d_ring_runs = nullptr;
    int4 *d_ring_runs_flip = nullptr;
    int2 *d_ring_kern = nullptr;
    float *d_ring_scale = nullptr;
    bool uploaded = false;
    std::mutex upload_mutex;
    // The runs and the ring runs as byte records for an LDS tile of row pitch
    // lp (lpa::DevKset::runs_b, lpa::DevRings::runs_b), one pair of tables
    // per pitch that a tile configuration or the adjoint uses (built at upload).
    struct ChainRuns {
        int lp;
        int4 *d_runs;
        int4 *d_ring_runs;
    };
    std::vector<ChainRuns> chain_runs;
    const ChainRuns *chain(int lp) const
    {
        for (const ChainRuns &t : chain_runs)
            if (t.lp == lp) return &t;
        return nullptr;
    }
    // The set for tiles of row pitch lp.
    lpa::DevKset dev(int lp) const
    {
        const ChainRuns *c = chain(lp);
        return lpa::DevKset{d_runs, d_weights, d_kern, c ? c->d_runs : nullptr, D, S};
    }
    // The ring runs as byte-offset pairs, one table per prefix row pitch that
    // a tile configuration or the adjoint uses (all built at upload).
    struct RingPairs {
        int lp;
        int4 *d_pairs;
        int2 *d_kern;
    };
    std::vector<RingPairs> ring_pairs;
    // Ring tables for a prefix table of row pitch lp and a float tile of row
    // pitch lp_tile (tables null where none was built).
    lpa::DevRings rings(int lp, int lp_tile) const
    {
        const RingPairs *rp = nullptr;
        for (const RingPairs &t : ring_pairs)
            if (t.lp == lp) rp = &t;
        const ChainRuns *c = chain(lp_tile);
        return lpa::DevRings{d_ring_runs,
                             d_ring_runs_flip,
                             d_ring_kern,
                             c ? c->d_ring_runs : nullptr,
                             rp ? rp->d_pairs : nullptr,
                             rp ? rp->d_kern : nullptr,
                             d_ring_scale,
                             d_ring_scale ? d_ring_scale + K : nullptr,
                             std::ldexp(1.0f, ring_frac_bits),
                             ring_range};
    }
    lpa::DevAdjoint adjoint() const
    {
        return lpa::DevAdjoint{lpa::DevKset{d_runs_flip, d_weights_flip, d_kern, nullptr, D, S},
                               d_tap_lds,
                               d_tap_kernel,
                               d_kern_w,
                               d_sumsq,
                               (int)weights.size()};
    }
};

namespace {

// Tile configurations used until one is set (lpa_kset_set_tile_config), the
// fastest on a 4096^2 image of the default set: 9 (1024 lanes within 64
// registers) for the fixed-point ring walk; 7 (512 lanes within 128
// registers) for the walks with more live values per pixel, which spill at 9.
constexpr int kDefaultShapeFixedPoint = 9, kDefaultShape = 7;

lpa_status hip_fail(hipError_t e, const char *what)
{
    return fail(e == hipErrorOutOfMemory ? LPA_ENOMEM : LPA_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                                                                              \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);                                                          \
    } while (0)

lpa::WalkMode walk_mode(const lpa_kset *ks, unsigned flags);

int shape_of(const lpa_kset *ks, unsigned flags)
{
    if (ks->shape >= 0) return ks->shape;
    return walk_mode(ks, flags) == lpa::kWalkRing ? kDefaultShapeFixedPoint : kDefaultShape;
}

// Row-run packing: for every kernel row with a nonzero, one run from the
// first to the last nonzero (interior zeros kept), offsets from the centre
// (rows/2, cols/2).  Follows build_sparse_kernels,
// src/v4_1_lpa_ici_2D_sparse_shared.cu:826-876, without its constant-memory
// caps (8192 weights / 1536 runs / 128 kernels).
void pack_runs(lpa_kset *ks, const float *dense, const int *offsets, const int *rows, const int *cols)
{
    int radius = 0;
    for (int k = 0; k < ks->K; ++k) {
        const float *w = dense + offsets[k];
        const int R = rows[k], C = cols[k];
        const int first_run = (int)ks->runs.size();
        int nz = 0;
        float ss = 0.0f;
        for (int i = 0; i < R * C; ++i) {
            ss += w[i] * w[i];  // float, dense row-major: compute_kernel_variances' sum
            nz += w[i] != 0.0f;
        }
        for (int r = 0; r < R; ++r) {
            int first = -1, last = -1;
            for (int c = 0; c < C; ++c)
                if (w[r * C + c] != 0.0f) {
                    if (first < 0) first = c;
                    last = c;
                }
            if (first < 0) continue;
            // A run longer than the kernels' longest unrolled body is cut into
            // consecutive pieces; the per-pixel tap order is unchanged.
            for (int c0 = first; c0 <= last; c0 += lpa::kMaxRun) {
                const int dr = r - R / 2, dc = c0 - C / 2, len = std::min(lpa::kMaxRun, last - c0 + 1);
                ks->runs.push_back(int4{dr, dc, len, (int)ks->weights.size()});
                for (int c = c0; c < c0 + len; ++c) ks->weights.push_back(w[r * C + c]);
                radius = std::max(radius, std::max(std::abs(dr), std::max(std::abs(dc), std::abs(dc + len - 1))));
            }
        }
        ks->kern.push_back(int2{first_run, (int)ks->runs.size() - first_run});
        const int first_w = first_run < (int)ks->runs.size() ? ks->runs[first_run].w : (int)ks->weights.size();
        int pieces = 0;
        for (int r = first_run; r < (int)ks->runs.size(); ++r)
            pieces += (ks->runs[r].z + lpa::kGradPiece - 1) / lpa::kGradPiece;
        ks->kern_w.push_back(int4{first_w, (int)ks->weights.size() - first_w, pieces, 0});
        ks->nnz.push_back(nz);
        ks->sumsq.push_back(ss);
        ks->num_nonzero += nz;
    }
    ks->radius = radius;
    // Adjoint tables: tap positions and owners, and the rotated runs.
    ks->weights_flip.resize(ks->weights.size());
    ks->tap_lds.resize(ks->weights.size());
    const int grad_pitch = lpa::lds_pitch(lpa::kBwdTileW, radius, lpa::kBwdPitchResidue);
    ks->tap_kernel.resize(ks->weights.size());
    for (int k = 0; k < ks->K; ++k) {
        const int2 kr = ks->kern[k];
        for (int r = kr.x; r < kr.x + kr.y; ++r) {
            const int4 run = ks->runs[r];
            ks->runs_flip.push_back(int4{-run.x, -(run.y + run.z - 1), run.z, run.w});
            for (int j = 0; j < run.z; ++j) {
                ks->weights_flip[run.w + j] = ks->weights[run.w + run.z - 1 - j];
                ks->tap_lds[run.w + j] = run.x * grad_pitch + run.y + j;
                ks->tap_kernel[run.w + j] = k;
            }
        }
    }
}

// Detect the ring form: for every direction, each scale's kernel has one
// weight value on its support and contains the previous scale's support.
// Then kernel (d,s) = w_s * indicator(support) and its correlation is w_s
// times a running sum over the rings support(d,s) \ support(d,s-1).
void build_rings(lpa_kset *ks, const float *dense, const int *offsets, const int *rows, const int *cols)
{
    std::vector<std::vector<std::pair<int, int>>> support(ks->K);
    for (int k = 0; k < ks->K; ++k) {
        const float *w = dense + offsets[k];
        const int R = rows[k], C = cols[k];
        float value = 0.0f;
        bool first = true;
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < C; ++c) {
                const float v = w[r * C + c];
                if (v == 0.0f) continue;
                if (first) {
                    value = v;
                    first = false;
                } else if (std::memcmp(&v, &value, sizeof v) != 0) {
                    return;  // weights not uniform
                }
                support[k].push_back({r - R / 2, c - C / 2});
            }
        ks->ring_scale.push_back(value);
    }
    for (int d = 0; d < ks->D; ++d)
        for (int s = 1; s < ks->S; ++s) {
            const auto &inner = support[d * ks->S + s - 1], &outer = support[d * ks->S + s];
            if (!std::includes(outer.begin(), outer.end(), inner.begin(), inner.end())) {
                ks->ring_scale.clear();
                return;  // supports not nested
            }
        }
    for (int d = 0; d < ks->D; ++d)
        for (int s = 0; s < ks->S; ++s) {
            const int k = d * ks->S + s;
            std::vector<std::pair<int, int>> ring;
            if (s == 0) {
                ring = support[k];
            } else {
                const auto &inner = support[k - 1];
                std::set_difference(support[k].begin(), support[k].end(), inner.begin(), inner.end(),
                                    std::back_inserter(ring));
            }
            const int first_run = (int)ks->ring_runs.size();
            for (size_t i = 0; i < ring.size();) {
                size_t j = i + 1;
                while (j < ring.size() && ring[j].first == ring[i].first && ring[j].second == ring[j - 1].second + 1 &&
                       (int)(j - i) < lpa::kMaxRun)
                    ++j;
                ks->ring_runs.push_back(int4{ring[i].first, ring[i].second, (int)(j - i), 0});
                i = j;
            }
            ks->ring_kern.push_back(int2{first_run, (int)ks->ring_runs.size() - first_run});
            ks->ring_taps += (int)ring.size();
        }
    for (const int4 &r : ks->ring_runs) ks->ring_runs_flip.push_back(int4{-r.x, -(r.y + r.z - 1), r.z, 0});
    // Fixed point for ring sums: with n_max the largest support, F is the
    // largest power making n_max * 2 * 2^F < 2^31, and the range bound r
    // keeps every sum exact: |rint(v 2^F)| < r 2^F + 1/2 for |v| < r, so a
    // support sum stays below n_max (r 2^F + 1/2) < 2^31.
    size_t nmax = 1;
    for (const auto &sp : support) nmax = std::max(nmax, sp.size());
    const double lim = 2147483647.0;
    int F = (int)std::floor(std::log2(lim / (2.0 * (double)nmax)));
    F = std::max(0, std::min(F, 30));
    ks->ring_frac_bits = F;
    ks->ring_range = (float)(lim * (1.0 - 1.0 / 4096.0) / ((double)nmax * std::ldexp(1.0, F)));
    ks->has_rings = true;
}

// Per-image [3][K] table: variance, threshold gamma*sqrt(var) and 1/var, in
// float on the host exactly as the reference forms them
// (src/v0_0_lpa_ici_2D_naive.cpp:684-699 and :887, :918-919).
void image_table(const lpa_kset *ks, const lpa_image &img, float *tab)
{
    const int K = ks->K;
    const float s2 = img.sigma * img.sigma;
    for (int k = 0; k < K; ++k) {
        const float var = img.variances ? img.variances[k] : s2 * ks->sumsq[k];
        tab[k] = var;
        tab[K + k] = img.gamma * std::sqrt(var);
        tab[2 * K + k] = 1.0f / var;
    }
}

lpa_status check_images(const lpa_kset *ks, const lpa_image *imgs, int n)
{
    if (!ks) return fail(LPA_EINVAL, "null kernel set");
    if (n <= 0 || !imgs) return fail(LPA_EINVAL, "no images");
    for (int i = 0; i < n; ++i) {
        const lpa_image &m = imgs[i];
        const std::string at = "image " + std::to_string(i) + ": ";
        if (!m.x || !m.y) return fail(LPA_EINVAL, at + "null x or y");
        if (m.width <= 0 || m.height <= 0) return fail(LPA_EINVAL, at + "width and height must be positive");
        if (m.width > 65535 || m.height > 65535 || (long long)m.width * m.height > (1LL << 31) - 1)
            return fail(LPA_ETOOBIG, at + "image too large");
        if (m.pitch != 0 && m.pitch < m.width) return fail(LPA_EINVAL, at + "pitch smaller than width");
        if (!m.variances && !(m.sigma > 0.0f)) return fail(LPA_EINVAL, at + "sigma must be positive");
        if (!(m.gamma > 0.0f)) return fail(LPA_EINVAL, at + "gamma must be positive");
    }
    return LPA_OK;
}

int tiles_x_of(int shape, int W) { return (W + lpa::kShapes[shape].tw - 1) / lpa::kShapes[shape].tw; }
int tiles_y_of(int shape, int H) { return (H + lpa::kShapes[shape].th - 1) / lpa::kShapes[shape].th; }

int num_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

// Device copies of the layout, made on the first call that needs them (on
// the device current at that call).  Thread-safe.
lpa_status ensure_uploaded(lpa_kset *ks)
{
    std::lock_guard<std::mutex> lock(ks->upload_mutex);
    if (ks->uploaded) return LPA_OK;
    const int K = ks->K;
    hipError_t e = hipGetDevice(&ks->device);
    const size_t nr = std::max<size_t>(1, ks->runs.size()), nw = std::max<size_t>(1, ks->weights.size());
    if (e == hipSuccess) e = hipMalloc(&ks->d_runs, sizeof(int4) * nr);
    // Weights get kMaxRun floats of slack: the chain body loads kMaxRun
    // weights from a run's first, whatever its length.
    if (e == hipSuccess) e = hipMalloc(&ks->d_weights, sizeof(float) * (nw + lpa::kMaxRun));
    if (e == hipSuccess) e = hipMalloc(&ks->d_kern, sizeof(int2) * K);
    if (e == hipSuccess) e = hipMemset(ks->d_weights, 0, sizeof(float) * (nw + lpa::kMaxRun));
    if (e == hipSuccess && !ks->runs.empty())
        e = hipMemcpy(ks->d_runs, ks->runs.data(), sizeof(int4) * ks->runs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !ks->weights.empty())
        e = hipMemcpy(ks->d_weights, ks->weights.data(), sizeof(float) * ks->weights.size(),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ks->d_kern, ks->kern.data(), sizeof(int2) * K, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&ks->d_runs_flip, sizeof(int4) * nr);
    if (e == hipSuccess) e = hipMalloc(&ks->d_weights_flip, sizeof(float) * (nw + lpa::kMaxRun));
    if (e == hipSuccess) e = hipMalloc(&ks->d_tap_lds, sizeof(int) * nw);
    if (e == hipSuccess) e = hipMalloc(&ks->d_tap_kernel, sizeof(int) * nw);
    if (e == hipSuccess) e = hipMalloc(&ks->d_kern_w, sizeof(int4) * K);
    if (e == hipSuccess) e = hipMalloc(&ks->d_sumsq, sizeof(float) * K);
    if (e == hipSuccess) e = hipMemset(ks->d_weights_flip, 0, sizeof(float) * (nw + lpa::kMaxRun));
    if (e == hipSuccess && !ks->runs.empty())
        e = hipMemcpy(ks->d_runs_flip, ks->runs_flip.data(), sizeof(int4) * ks->runs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !ks->weights.empty()) {
        const size_t n = ks->weights.size();
        e = hipMemcpy(ks->d_weights_flip, ks->weights_flip.data(), sizeof(float) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ks->d_tap_lds, ks->tap_lds.data(), sizeof(int) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(ks->d_tap_kernel, ks->tap_kernel.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(ks->d_kern_w, ks->kern_w.data(), sizeof(int4) * K, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ks->d_sumsq, ks->sumsq.data(), sizeof(float) * K, hipMemcpyHostToDevice);
    if (e == hipSuccess && ks->has_rings) {
        const size_t nrr = std::max<size_t>(1, ks->ring_runs.size());
        e = hipMalloc(&ks->d_ring_runs, sizeof(int4) * nrr);
        if (e == hipSuccess) e = hipMalloc(&ks->d_ring_runs_flip, sizeof(int4) * nrr);
        if (e == hipSuccess && !ks->ring_runs_flip.empty())
            e = hipMemcpy(ks->d_ring_runs_flip, ks->ring_runs_flip.data(), sizeof(int4) * ks->ring_runs_flip.size(),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&ks->d_ring_kern, sizeof(int2) * K);
        // [2][K]: w_s, then w_s * 2^-F (the fixed-point walk's factor)
        if (e == hipSuccess) e = hipMalloc(&ks->d_ring_scale, sizeof(float) * 2 * K);
        if (e == hipSuccess && !ks->ring_runs.empty())
            e = hipMemcpy(ks->d_ring_runs, ks->ring_runs.data(), sizeof(int4) * ks->ring_runs.size(),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(ks->d_ring_kern, ks->ring_kern.data(), sizeof(int2) * K, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            std::vector<float> scales(ks->ring_scale);
            for (int k = 0; k < K; ++k) scales.push_back(std::ldexp(ks->ring_scale[k], -ks->ring_frac_bits));
            e = hipMemcpy(ks->d_ring_scale, scales.data(), sizeof(float) * 2 * K, hipMemcpyHostToDevice);
        }
        // Byte-offset pairs for every prefix row pitch in use: the forward's
        // tile configurations and the adjoint's 64-wide tiles.
        std::vector<int> pitches{lpa::prefix_pitch(lpa::kBwdTileW, ks->radius)};
        for (int i = 0; i < lpa::kNumShapes; ++i) {
            const int lp = lpa::prefix_pitch(lpa::kShapes[i].tw, ks->radius, lpa::kShapes[i].pm);
            if (std::find(pitches.begin(), pitches.end(), lp) == pitches.end()) pitches.push_back(lp);
        }
        for (int lp : pitches) {
            if (e != hipSuccess) break;
            std::vector<int4> pairs;
            std::vector<int2> pkern;
            for (int k = 0; k < K; ++k) {
                const int2 kr = ks->ring_kern[k];
                pkern.push_back(int2{(int)pairs.size(), kr.y});
                for (int r = kr.x; r < kr.x + kr.y; r += 2) {
                    const int4 u = ks->ring_runs[r];
                    const int4 v = r + 1 < kr.x + kr.y ? ks->ring_runs[r + 1] : u;
                    const int au = 4 * (u.x * lp + u.y), av = 4 * (v.x * lp + v.y);
                    pairs.push_back(int4{au, au + 4 * u.z, av, av + 4 * v.z});
                }
            }
            if (pairs.empty()) pairs.push_back(int4{0, 0, 0, 0});
            lpa_kset::RingPairs t{lp, nullptr, nullptr};
            e = hipMalloc(&t.d_pairs, sizeof(int4) * pairs.size());
            if (e == hipSuccess) e = hipMalloc(&t.d_kern, sizeof(int2) * K);
            if (e == hipSuccess)
                e = hipMemcpy(t.d_pairs, pairs.data(), sizeof(int4) * pairs.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(t.d_kern, pkern.data(), sizeof(int2) * K, hipMemcpyHostToDevice);
            ks->ring_pairs.push_back(t);
        }
    }
    // Byte records of the runs and ring runs for every tile row pitch in use:
    // the forward's tile configurations (float tile and prefix table share
    // the pitch) and the adjoint's float tile.
    std::vector<int> tile_pitches{lpa::lds_pitch(lpa::kBwdTileW, ks->radius, lpa::kBwdPitchResidue)};
    for (int i = 0; i < lpa::kNumShapes; ++i) {
        const int lp = lpa::prefix_pitch(lpa::kShapes[i].tw, ks->radius, lpa::kShapes[i].pm);
        if (std::find(tile_pitches.begin(), tile_pitches.end(), lp) == tile_pitches.end()) tile_pitches.push_back(lp);
    }
    for (int lp : tile_pitches) {
        if (e != hipSuccess) break;
        auto records = [lp](const std::vector<int4> &runs, bool weighted) {
            std::vector<int4> out;
            for (const int4 &r : runs) out.push_back(int4{4 * (r.x * lp + r.y), r.z, weighted ? 4 * r.w : 0, 0});
            if (out.empty()) out.push_back(int4{0, 1, 0, 0});
            return out;
        };
        lpa_kset::ChainRuns t{lp, nullptr, nullptr};
        const std::vector<int4> rb = records(ks->runs, true);
        e = hipMalloc(&t.d_runs, sizeof(int4) * rb.size());
        if (e == hipSuccess) e = hipMemcpy(t.d_runs, rb.data(), sizeof(int4) * rb.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess && ks->has_rings) {
            const std::vector<int4> rr = records(ks->ring_runs, false);
            e = hipMalloc(&t.d_ring_runs, sizeof(int4) * rr.size());
            if (e == hipSuccess)
                e = hipMemcpy(t.d_ring_runs, rr.data(), sizeof(int4) * rr.size(), hipMemcpyHostToDevice);
        }
        ks->chain_runs.push_back(t);
    }
    if (e != hipSuccess) return hip_fail(e, "kernel set upload");
    ks->uploaded = true;
    return LPA_OK;
}

}  // namespace

// Device state of a batch: per-image descriptors and tables, the tile queue
// and its counter, in one allocation.
struct lpa_plan {
    lpa_kset *ks = nullptr;
    unsigned flags = 0;
    int shape = 0;
    int n = 0;
    int num_tiles = 0;
    int grid = 0;
    std::vector<lpa::ImgDesc> host_desc;  // device pointers, for the plain form
    void *d_mem = nullptr;
    lpa::ImgDesc *d_desc = nullptr;
    uint2 *d_tiles = nullptr;
    int *d_counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Where a plan's device state lives: a stream-ordered allocation freed after
// the launch (lpa_ici_forward), an allocation the plan owns (lpa_plan_create),
// or a caller-provided buffer of `cap` bytes that nothing allocates or frees
// (lpa_ici_forward_noalloc).
struct PlanMemory {
    enum Kind { kAsync, kOwned, kBuffer } kind;
    void *buffer;
    size_t cap;
};

// Lay out one plan allocation: [counter | descriptors | tables | tiles].
lpa_status plan_build(lpa_kset *ks, const lpa_image *imgs, int n, unsigned flags, PlanMemory mem, hipStream_t stream,
                      lpa_plan *p)
{
    p->ks = ks;
    p->flags = flags;
    p->shape = shape_of(ks, flags);
    p->n = n;
    const int K = ks->K;
    std::vector<uint2> tiles;
    for (int i = 0; i < n; ++i) {
        const int tx = tiles_x_of(p->shape, imgs[i].width), ty = tiles_y_of(p->shape, imgs[i].height);
        for (int y = 0; y < ty; ++y)
            for (int x = 0; x < tx; ++x) tiles.push_back(uint2{(unsigned)i, ((unsigned)y << 16) | (unsigned)x});
    }
    p->num_tiles = (int)tiles.size();
    const size_t off_desc = 256;
    const size_t off_tab = off_desc + ((sizeof(lpa::ImgDesc) * n + 255) / 256) * 256;
    const size_t off_tiles = off_tab + ((sizeof(float) * 3 * K * n + 255) / 256) * 256;
    const size_t bytes = off_tiles + sizeof(uint2) * tiles.size();
    unsigned char *base = nullptr;
    if (mem.kind == PlanMemory::kBuffer) {
        if (bytes > mem.cap)
            return fail(LPA_ETOOBIG, "the batch needs " + std::to_string(bytes) + " bytes of workspace (descriptors, " +
                                         "tables, tile queue); the static workspace holds " +
                                         std::to_string(mem.cap));
        base = (unsigned char *)mem.buffer;
    } else {
        if (mem.kind == PlanMemory::kAsync) {
            HIP_TRY(hipMallocAsync(&p->d_mem, bytes, stream));
        } else {
            HIP_TRY(hipMalloc(&p->d_mem, bytes));
        }
        base = (unsigned char *)p->d_mem;
    }
    std::vector<unsigned char> host(bytes, 0);
    p->d_counter = (int *)base;
    p->d_desc = (lpa::ImgDesc *)(base + off_desc);
    p->d_tiles = (uint2 *)(base + off_tiles);
    float *h_tab = (float *)(host.data() + off_tab);
    const float *d_tab = (const float *)(base + off_tab);
    p->host_desc.resize(n);
    for (int i = 0; i < n; ++i) {
        const lpa_image &m = imgs[i];
        image_table(ks, m, h_tab + (size_t)3 * K * i);
        lpa::ImgDesc &d = p->host_desc[i];
        d.x = m.x;
        d.y = m.y;
        d.map = m.scale_map;
        d.est = m.estimates;
        d.tab = d_tab + (size_t)3 * K * i;
        d.W = m.width;
        d.H = m.height;
        d.pitch = m.pitch ? m.pitch : m.width;
        d.tiles_x = tiles_x_of(p->shape, m.width);
    }
    std::memcpy(host.data() + off_desc, p->host_desc.data(), sizeof(lpa::ImgDesc) * n);
    std::memcpy(host.data() + off_tiles, tiles.data(), sizeof(uint2) * tiles.size());
    HIP_TRY(hipMemcpyAsync(base, host.data(), bytes, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));  // `host` goes out of scope on return
    const int bpc = lpa::blocks_per_cu(p->shape, ks->radius, ks->K);
    p->grid = std::max(1, std::min(p->num_tiles, num_cus() * bpc));
    return LPA_OK;
}

// Device address of `scratch` on the current device, looked up on the first
// call per device and kept (each device holds its own copy of the array).
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void *scratch_ptr[kMaxDevices] = {};

lpa_status scratch_address(void **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LPA_EINVAL, "device index " + std::to_string(dev) + " out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(LPA_EHIP, "hipGetSymbolAddress returned a null address for the static workspace");
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return LPA_OK;
}

lpa::WalkMode walk_mode(const lpa_kset *ks, unsigned flags)
{
    if (flags & LPA_FWD_EXACT) return lpa::kWalkExact;
    if (ks->has_rings && !(flags & LPA_FWD_WEIGHTED))
        return (flags & LPA_FWD_RING_FLOAT) ? lpa::kWalkRingFloat : lpa::kWalkRing;
    return lpa::kWalkFma;
}

lpa_status plan_launch(lpa_plan *p, hipStream_t stream)
{
    const lpa::WalkMode mode = walk_mode(p->ks, p->flags);
    const lpa::TileShape &sh = lpa::kShapes[p->shape];
    const int lp = lpa::prefix_pitch(sh.tw, p->ks->radius, sh.pm);  // float tile and prefix table alike
    const lpa::DevKset dk = p->ks->dev(lp);
    const lpa::DevRings dr = p->ks->rings(lp, lp);
    if (p->flags & LPA_FWD_PLAIN) {
        for (int i = 0; i < p->n; ++i) {
            hipError_t e = lpa::launch_forward_plain(p->shape, mode, dk, dr, p->host_desc[i], p->ks->radius, stream);
            if (e != hipSuccess) return hip_fail(e, "forward launch");
        }
    } else {
        hipError_t e = lpa::launch_forward_batched(p->shape, mode, dk, dr, p->d_desc, p->d_tiles, p->num_tiles,
                                                   p->d_counter, p->ks->radius, p->grid, stream);
        if (e != hipSuccess) return hip_fail(e, "batched forward launch");
    }
    return LPA_OK;
}

}  // namespace

extern "C" {

lpa_status lpa_kset_create(int num_dirs, int num_scales, const float *dense, const int *offsets, const int *rows,
                           const int *cols, lpa_kset_t *out)
{
    if (!out) return fail(LPA_EINVAL, "null output handle");
    *out = nullptr;
    if (num_dirs <= 0 || num_scales <= 0 || num_scales > 255 || !dense || !offsets || !rows || !cols)
        return fail(LPA_EINVAL, "bad kernel set arguments");
    const int K = num_dirs * num_scales;
    for (int k = 0; k < K; ++k) {
        if (rows[k] <= 0 || cols[k] <= 0 || offsets[k] < 0)
            return fail(LPA_EINVAL, "kernel " + std::to_string(k) + " has a bad size or offset");
        if (rows[k] / 2 > lpa::kMaxRadius || cols[k] / 2 > lpa::kMaxRadius)
            return fail(LPA_ETOOBIG, "kernel " + std::to_string(k) + " is " + std::to_string(rows[k]) + "x" +
                                         std::to_string(cols[k]) + ", max supported is " +
                                         std::to_string(2 * lpa::kMaxRadius + 1) + "x" +
                                         std::to_string(2 * lpa::kMaxRadius + 1));
    }
    lpa_kset *ks = new (std::nothrow) lpa_kset;
    if (!ks) return fail(LPA_ENOMEM, "out of host memory");
    ks->D = num_dirs;
    ks->S = num_scales;
    ks->K = K;
    pack_runs(ks, dense, offsets, rows, cols);
    build_rings(ks, dense, offsets, rows, cols);
    *out = ks;
    lpa::clear_error();
    return LPA_OK;
}

void lpa_kset_destroy(lpa_kset_t ks)
{
    if (!ks) return;
    if (ks->d_runs) (void)hipFree(ks->d_runs);
    if (ks->d_weights) (void)hipFree(ks->d_weights);
    if (ks->d_kern) (void)hipFree(ks->d_kern);
    for (void *p : {(void *)ks->d_runs_flip, (void *)ks->d_weights_flip, (void *)ks->d_tap_lds,
                    (void *)ks->d_tap_kernel, (void *)ks->d_kern_w, (void *)ks->d_sumsq, (void *)ks->d_ring_runs,
                    (void *)ks->d_ring_runs_flip, (void *)ks->d_ring_kern, (void *)ks->d_ring_scale})
        if (p) (void)hipFree(p);
    for (const auto &t : ks->ring_pairs) {
        if (t.d_pairs) (void)hipFree(t.d_pairs);
        if (t.d_kern) (void)hipFree(t.d_kern);
    }
    for (const auto &t : ks->chain_runs) {
        if (t.d_runs) (void)hipFree(t.d_runs);
        if (t.d_ring_runs) (void)hipFree(t.d_ring_runs);
    }
    delete ks;
}

lpa_status lpa_kset_get_info(lpa_kset_t ks, lpa_kset_info *info)
{
    if (!ks || !info) return fail(LPA_EINVAL, "null argument");
    info->num_dirs = ks->D;
    info->num_scales = ks->S;
    info->num_runs = (int)ks->runs.size();
    info->num_weights = (int)ks->weights.size();
    info->num_nonzero = ks->num_nonzero;
    info->max_radius = ks->radius;
    info->ring_form = ks->has_rings ? 1 : 0;
    info->num_ring_taps = ks->ring_taps;
    info->ring_frac_bits = ks->ring_frac_bits;
    info->ring_range = ks->ring_range;
    return LPA_OK;
}

lpa_status lpa_kset_get_runs(lpa_kset_t ks, int *run_row, int *run_col, int *run_len, int *run_woff,
                             int *kernel_first_run, int *kernel_num_runs, float *weights, int *nnz, float *sumsq)
{
    if (!ks) return fail(LPA_EINVAL, "null kernel set");
    for (size_t r = 0; r < ks->runs.size(); ++r) {
        if (run_row) run_row[r] = ks->runs[r].x;
        if (run_col) run_col[r] = ks->runs[r].y;
        if (run_len) run_len[r] = ks->runs[r].z;
        if (run_woff) run_woff[r] = ks->runs[r].w;
    }
    for (int k = 0; k < ks->K; ++k) {
        if (kernel_first_run) kernel_first_run[k] = ks->kern[k].x;
        if (kernel_num_runs) kernel_num_runs[k] = ks->kern[k].y;
        if (nnz) nnz[k] = ks->nnz[k];
        if (sumsq) sumsq[k] = ks->sumsq[k];
    }
    if (weights && !ks->weights.empty()) std::memcpy(weights, ks->weights.data(), sizeof(float) * ks->weights.size());
    return LPA_OK;
}

int lpa_num_tile_configs(void) { return lpa::kNumShapes; }

lpa_status lpa_tile_config_get(int index, lpa_tile_config *cfg)
{
    if (!cfg || index < 0 || index >= lpa::kNumShapes) return fail(LPA_EINVAL, "bad tile configuration index");
    cfg->tile_w = lpa::kShapes[index].tw;
    cfg->tile_h = lpa::kShapes[index].th;
    cfg->px_per_lane = lpa::kShapes[index].px;
    cfg->body = lpa::kShapes[index].body;
    cfg->halo = lpa::kMaxRadius;
    int devices = 0;
    cfg->blocks_per_cu =
        hipGetDeviceCount(&devices) == hipSuccess && devices > 0 ? lpa::blocks_per_cu(index, lpa::kMaxRadius, 0) : 0;
    return LPA_OK;
}

lpa_status lpa_kset_set_tile_config(lpa_kset_t ks, int index)
{
    if (!ks || index < -1 || index >= lpa::kNumShapes) return fail(LPA_EINVAL, "bad tile configuration index");
    ks->shape = index;
    return LPA_OK;
}

int lpa_kset_get_tile_config(lpa_kset_t ks) { return ks ? shape_of(ks, 0) : -1; }

lpa_status lpa_ici_forward(lpa_kset_t ks, const lpa_image *imgs, int n, unsigned flags, void *stream)
{
    lpa_status st = check_images(ks, imgs, n);
    if (st != LPA_OK) return st;
    if ((st = ensure_uploaded(ks)) != LPA_OK) return st;
    if (flags & LPA_FWD_GRAPH) return fail(LPA_EINVAL, "LPA_FWD_GRAPH applies to plans only");
    hipStream_t s = (hipStream_t)stream;
    lpa_plan p;
    st = plan_build(ks, imgs, n, flags, PlanMemory{PlanMemory::kAsync, nullptr, 0}, s, &p);
    if (st == LPA_OK) st = plan_launch(&p, s);
    if (p.d_mem) {
        hipError_t e = hipFreeAsync(p.d_mem, s);
        if (st == LPA_OK && e != hipSuccess) st = hip_fail(e, "hipFreeAsync");
    }
    if (st == LPA_OK) lpa::clear_error();
    return st;
}

lpa_status lpa_ici_forward_noalloc(lpa_kset_t ks, const lpa_image *imgs, int n, unsigned flags, void *stream)
{
    lpa_status st = check_images(ks, imgs, n);
    if (st != LPA_OK) return st;
    if (flags & LPA_FWD_GRAPH) return fail(LPA_EINVAL, "LPA_FWD_GRAPH applies to plans only");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    {
        // The set's device tables are allocated on its first use; this entry
        // point allocates nothing, so the set must be resident already.
        std::lock_guard<std::mutex> lock(ks->upload_mutex);
        if (!ks->uploaded)
            return fail(LPA_EINVAL, "the kernel set is not resident on a device yet (its first lpa_ici_forward, "
                                    "lpa_plan_create, lpa_ici_estimates or lpa_ici_backward call uploads it)");
        if (ks->device != dev)
            return fail(LPA_EINVAL, "the kernel set is resident on device " + std::to_string(ks->device) +
                                        ", not on the current device " + std::to_string(dev));
    }
    void *buf = nullptr;
    if ((st = scratch_address(&buf)) != LPA_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    lpa_plan p;
    st = plan_build(ks, imgs, n, flags, PlanMemory{PlanMemory::kBuffer, buf, kStaticWorkspaceBytes}, s, &p);
    if (st == LPA_OK) st = plan_launch(&p, s);
    if (st == LPA_OK) lpa::clear_error();
    return st;
}

lpa_status lpa_ici_estimates(lpa_kset_t ks, const lpa_image *imgs, float *const *est, int n, unsigned flags,
                             void *stream)
{
    lpa_status st = check_images(ks, imgs, n);
    if (st != LPA_OK) return st;
    if ((st = ensure_uploaded(ks)) != LPA_OK) return st;
    if (flags & LPA_FWD_GRAPH) return fail(LPA_EINVAL, "LPA_FWD_GRAPH applies to plans only");
    if (!est) return fail(LPA_EINVAL, "null est array");
    for (int i = 0; i < n; ++i) {
        if (!imgs[i].scale_map)
            return fail(LPA_EINVAL, "image " + std::to_string(i) + ": the scale map of a forward is required");
        if (!est[i]) return fail(LPA_EINVAL, "image " + std::to_string(i) + ": null est");
    }
    hipStream_t s = (hipStream_t)stream;
    const lpa::WalkMode mode = walk_mode(ks, flags);
    const int SH = mode == lpa::kWalkRing ? shape_of(ks, flags) : lpa::kEstimateShape;
    const int est_lp = lpa::prefix_pitch(lpa::kShapes[SH].tw, ks->radius, lpa::kShapes[SH].pm);
    std::vector<uint2> tiles;
    for (int i = 0; i < n; ++i) {
        const int tx = tiles_x_of(SH, imgs[i].width), ty = tiles_y_of(SH, imgs[i].height);
        for (int y = 0; y < ty; ++y)
            for (int x = 0; x < tx; ++x) tiles.push_back(uint2{(unsigned)i, ((unsigned)y << 16) | (unsigned)x});
    }
    const int bpc = lpa::blocks_per_cu(SH, ks->radius, ks->K);
    const int grid = std::max(1, std::min((int)tiles.size(), num_cus() * bpc));
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t off_desc = 256;
    const size_t off_tiles = off_desc + up(sizeof(lpa::EstDesc) * n);
    const size_t bytes = off_tiles + sizeof(uint2) * tiles.size();
    std::vector<unsigned char> host(bytes, 0);
    std::vector<lpa::EstDesc> desc(n);
    for (int i = 0; i < n; ++i) {
        const lpa_image &m = imgs[i];
        lpa::EstDesc &e = desc[i];
        e.img.x = m.x;
        e.img.y = m.y;
        e.img.map = m.scale_map;
        e.img.est = nullptr;
        e.img.tab = nullptr;
        e.img.W = m.width;
        e.img.H = m.height;
        e.img.pitch = m.pitch ? m.pitch : m.width;
        e.img.tiles_x = tiles_x_of(SH, m.width);
        e.est = est[i];
    }
    std::memcpy(host.data() + off_desc, desc.data(), sizeof(lpa::EstDesc) * n);
    std::memcpy(host.data() + off_tiles, tiles.data(), sizeof(uint2) * tiles.size());
    void *mem = nullptr;
    HIP_TRY(hipMallocAsync(&mem, bytes, s));
    unsigned char *base = (unsigned char *)mem;
    hipError_t e = hipMemcpyAsync(mem, host.data(), bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `host` is freed on return
    if (e == hipSuccess)
        e = lpa::launch_estimates(mode, SH, ks->dev(est_lp), ks->has_rings ? ks->rings(est_lp, est_lp) : lpa::DevRings{},
                                  (const lpa::EstDesc *)(base + off_desc), (const uint2 *)(base + off_tiles),
                                  (int)tiles.size(), (int *)base, ks->radius, grid, s);
    hipError_t ef = hipFreeAsync(mem, s);
    if (e != hipSuccess) return hip_fail(e, "estimates");
    if (ef != hipSuccess) return hip_fail(ef, "hipFreeAsync");
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_ici_backward(lpa_kset_t ks, const lpa_image *imgs, const float *const *dy, float *const *dx,
                            float *dkernel, int n, void *stream)
{
    lpa_status st = check_images(ks, imgs, n);
    if (st != LPA_OK) return st;
    if ((st = ensure_uploaded(ks)) != LPA_OK) return st;
    if (!dy || !dx) return fail(LPA_EINVAL, "null dy or dx array");
    if (ks->S > 16) return fail(LPA_ETOOBIG, "the adjoint supports at most 16 scales");
    const int bwd_lp = lpa::lds_pitch(lpa::kBwdTileW, ks->radius, lpa::kBwdPitchResidue);  // pass 1's float tile
    for (int i = 0; i < n; ++i) {
        if (!imgs[i].scale_map) return fail(LPA_EINVAL, "image " + std::to_string(i) + ": the adjoint needs the scale map");
        if (imgs[i].variances)
            return fail(LPA_EINVAL, "image " + std::to_string(i) +
                                        ": the adjoint differentiates sigma^2*sum(w^2); explicit variances are not allowed");
        if (!dy[i] || !dx[i]) return fail(LPA_EINVAL, "image " + std::to_string(i) + ": null dy or dx");
    }
    hipStream_t s = (hipStream_t)stream;
    const int K = ks->K, NW = (int)ks->weights.size();
    // Tiles and tables as for the batched forward, shape 0.
    std::vector<uint2> tiles, dx_tiles;
    auto add_tiles = [&](std::vector<uint2> &v, int tw, int th) {
        for (int i = 0; i < n; ++i) {
            const int tx = (imgs[i].width + tw - 1) / tw, ty = (imgs[i].height + th - 1) / th;
            for (int y = 0; y < ty; ++y)
                for (int x = 0; x < tx; ++x) v.push_back(uint2{(unsigned)i, ((unsigned)y << 16) | (unsigned)x});
        }
    };
    add_tiles(tiles, lpa::kBwdTileW, lpa::kBwdTileH);
    add_tiles(dx_tiles, lpa::kDxTileW, ks->has_rings ? lpa::kDxTileRowsRing : lpa::kDxTileRowsWeighted);
    int bp = 1, bd = 1;
    lpa::backward_occupancy(ks->radius, ks->has_rings, K, &bp, &bd);
    const int cus = num_cus();
    const int grid_prep = std::max(1, std::min((int)tiles.size(), cus * bp));
    const int grid_dx = std::max(1, std::min((int)dx_tiles.size(), cus * bd));
    size_t qpx = 0;
    for (int i = 0; i < n; ++i) qpx += (size_t)imgs[i].width * imgs[i].height;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t off_desc = 256;
    const size_t off_tab = off_desc + up(sizeof(lpa::BwdDesc) * n);
    const size_t off_tiles = off_tab + up(sizeof(float) * 3 * K * n);
    const size_t off_dxt = off_tiles + up(sizeof(uint2) * tiles.size());
    const size_t off_q = off_dxt + up(sizeof(uint2) * dx_tiles.size());
    const size_t off_part = off_q + up(sizeof(float) * qpx);
    const size_t bytes = off_part + sizeof(float) * ((size_t)grid_prep * (NW + K) + 1 + 2 * (size_t)(NW + K));
    const size_t host_bytes = off_q;
    std::vector<unsigned char> host(host_bytes, 0);
    void *mem = nullptr;
    HIP_TRY(hipMallocAsync(&mem, bytes, s));
    unsigned char *base = (unsigned char *)mem;
    std::vector<lpa::BwdDesc> desc(n);
    float *h_tab = (float *)(host.data() + off_tab);
    size_t qoff = 0;
    for (int i = 0; i < n; ++i) {
        const lpa_image &m = imgs[i];
        image_table(ks, m, h_tab + (size_t)3 * K * i);
        lpa::BwdDesc &b = desc[i];
        b.img.x = m.x;
        b.img.y = m.y;
        b.img.map = m.scale_map;
        b.img.est = m.estimates;
        b.img.tab = (const float *)(base + off_tab) + (size_t)3 * K * i;
        b.img.W = m.width;
        b.img.H = m.height;
        b.img.pitch = m.pitch ? m.pitch : m.width;
        b.img.tiles_x = tiles_x_of(0, m.width);
        b.g = dy[i];
        b.dx = dx[i];
        b.q = (float *)(base + off_q) + qoff;
        qoff += (size_t)m.width * m.height;
    }
    std::memcpy(host.data() + off_desc, desc.data(), sizeof(lpa::BwdDesc) * n);
    std::memcpy(host.data() + off_tiles, tiles.data(), sizeof(uint2) * tiles.size());
    std::memcpy(host.data() + off_dxt, dx_tiles.data(), sizeof(uint2) * dx_tiles.size());
    hipError_t e = hipMemcpyAsync(mem, host.data(), host_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `host` is freed on return
    if (e == hipSuccess)
        e = lpa::launch_backward(ks->dev(bwd_lp),
                                 ks->has_rings ? ks->rings(lpa::prefix_pitch(lpa::kBwdTileW, ks->radius), bwd_lp)
                                               : lpa::DevRings{},
                                 ks->adjoint(), (const lpa::BwdDesc *)(base + off_desc),
                                 (const uint2 *)(base + off_tiles), (int)tiles.size(),
                                 (const uint2 *)(base + off_dxt), (int)dx_tiles.size(), (int *)base, ks->radius,
                                 grid_prep, grid_dx, (float *)(base + off_part), dkernel, s);
    hipError_t ef = hipFreeAsync(mem, s);
    if (e != hipSuccess) return hip_fail(e, "backward");
    if (ef != hipSuccess) return hip_fail(ef, "hipFreeAsync");
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_workspace_peak(unsigned long long *bytes, int reset)
{
    if (!bytes) return fail(LPA_EINVAL, "null output");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipMemPool_t pool = nullptr;
    HIP_TRY(hipDeviceGetDefaultMemPool(&pool, dev));
    uint64_t high = 0;
    HIP_TRY(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemHigh, &high));
    *bytes = (unsigned long long)high;
    if (reset) {
        uint64_t zero = 0;
        HIP_TRY(hipMemPoolSetAttribute(pool, hipMemPoolAttrUsedMemHigh, &zero));
    }
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_plan_create(lpa_kset_t ks, const lpa_image *imgs, int n, unsigned flags, lpa_plan_t *out)
{
    if (!out) return fail(LPA_EINVAL, "null output handle");
    *out = nullptr;
    lpa_status st = check_images(ks, imgs, n);
    if (st != LPA_OK) return st;
    if ((st = ensure_uploaded(ks)) != LPA_OK) return st;
    lpa_plan *p = new (std::nothrow) lpa_plan;
    if (!p) return fail(LPA_ENOMEM, "out of host memory");
    st = plan_build(ks, imgs, n, flags, PlanMemory{PlanMemory::kOwned, nullptr, 0}, nullptr, p);
    if (st != LPA_OK) {
        lpa_plan_destroy(p);
        return st;
    }
    *out = p;
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_plan_run(lpa_plan_t p, void *stream)
{
    if (!p) return fail(LPA_EINVAL, "null plan");
    hipStream_t s = (hipStream_t)stream;
    if (!(p->flags & LPA_FWD_GRAPH)) {
        lpa_status st = plan_launch(p, s);
        if (st == LPA_OK) lpa::clear_error();
        return st;
    }
    if (!s) return fail(LPA_EINVAL, "graph replay needs a non-default stream");
    if (!p->exec) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        lpa_status st = plan_launch(p, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st != LPA_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        p->graph = g;
        HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(p->exec, s));
    lpa::clear_error();
    return LPA_OK;
}

void lpa_plan_destroy(lpa_plan_t p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_mem) (void)hipFree(p->d_mem);
    delete p;
}

}  // extern "C"
