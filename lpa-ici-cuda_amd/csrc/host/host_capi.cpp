// C entry points of the host substrate (include/lpa_ici.h, last section).
// Plain C++, no HIP; compiled with -ffp-contract=off.
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../internal.hpp"
#include "lpa_ici.h"
#include "lpa_ici_host.hpp"

namespace lpa {
namespace {
thread_local std::string t_error;
}
lpa_status fail(lpa_status st, const std::string &msg)
{
    t_error = msg;
    return st;
}
void clear_error() { t_error.clear(); }
}  // namespace lpa

extern "C" {

const char *lpa_last_error(void) { return lpa::t_error.c_str(); }

lpa_status lpa_load_kernels(const char *path, int *num_dirs, int *num_scales, int *n_weights, float *weights,
                            int *offsets, int *rows, int *cols)
{
    if (!path || !num_dirs || !num_scales || !n_weights) return lpa::fail(LPA_EINVAL, "null argument");
    std::vector<float> w;
    std::vector<int> off, r, c;
    int D = 0, S = 0;
    std::string err = lpa::load_kernels(path, w, off, r, c, D, S);
    if (!err.empty()) return lpa::fail(err.rfind("Failed to open", 0) == 0 ? LPA_EIO : LPA_EINVAL, err);
    *num_dirs = D;
    *num_scales = S;
    *n_weights = (int)w.size();
    if (weights) std::memcpy(weights, w.data(), w.size() * sizeof(float));
    if (offsets) std::memcpy(offsets, off.data(), off.size() * sizeof(int));
    if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(int));
    if (cols) std::memcpy(cols, c.data(), c.size() * sizeof(int));
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_wedge_kernels(int num_dirs, int num_scales, const int *h, int *n_weights, float *weights,
                             int *offsets, int *rows, int *cols)
{
    if (!n_weights) return lpa::fail(LPA_EINVAL, "null argument");
    std::vector<float> w;
    std::vector<int> off, r, c;
    std::string err = lpa::wedge_kernels(num_dirs, num_scales, h, w, off, r, c);
    if (!err.empty()) return lpa::fail(LPA_EINVAL, err);
    *n_weights = (int)w.size();
    if (weights) std::memcpy(weights, w.data(), w.size() * sizeof(float));
    if (offsets) std::memcpy(offsets, off.data(), off.size() * sizeof(int));
    if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(int));
    if (cols) std::memcpy(cols, c.data(), c.size() * sizeof(int));
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_make_kernels(int num_dirs, int num_scales, const int *h, int order, int window, int *n_weights,
                            float *weights, int *offsets, int *rows, int *cols)
{
    if (!n_weights) return lpa::fail(LPA_EINVAL, "null argument");
    std::vector<float> w;
    std::vector<int> off, r, c;
    std::string err = lpa::lpa_kernels(num_dirs, num_scales, h, order, window, w, off, r, c);
    if (!err.empty()) return lpa::fail(LPA_EINVAL, err);
    *n_weights = (int)w.size();
    if (weights) std::memcpy(weights, w.data(), w.size() * sizeof(float));
    if (offsets) std::memcpy(offsets, off.data(), off.size() * sizeof(int));
    if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(int));
    if (cols) std::memcpy(cols, c.data(), c.size() * sizeof(int));
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_write_kernels(const char *path, int num_dirs, int num_scales, const float *weights,
                             const int *offsets, const int *rows, const int *cols)
{
    if (!path || !weights || !offsets || !rows || !cols || num_dirs < 1 || num_scales < 1)
        return lpa::fail(LPA_EINVAL, "null argument or empty kernel set");
    std::string err = lpa::write_kernels(path, num_dirs, num_scales, weights, offsets, rows, cols);
    if (!err.empty()) return lpa::fail(LPA_EIO, err);
    lpa::clear_error();
    return LPA_OK;
}

// var_k = (sigma*sigma) * sum(w^2): the square of sigma is rounded to float
// once, the sum runs in float over the dense kernel in row-major order.
void lpa_kernel_variances(int num_kernels, const float *weights, const int *offsets, const int *rows,
                          const int *cols, float sigma, float *variances)
{
    const float s2 = sigma * sigma;
    for (int k = 0; k < num_kernels; ++k) {
        const float *w = weights + offsets[k];
        float acc = 0.0f;
        for (int i = 0; i < rows[k] * cols[k]; ++i) acc += w[i] * w[i];
        variances[k] = s2 * acc;
    }
}

// One std::normal_distribution<float>(0, sigma) draw per pixel in raster
// order from std::mt19937(seed): the libstdc++ types fix every bit.
void lpa_make_noisy(const float *clean, float *noisy, long long n, float sigma, unsigned seed)
{
    std::mt19937 engine{seed};
    std::normal_distribution<float> gauss{0.0f, sigma};
    for (long long i = 0; i < n; ++i) noisy[i] = clean[i] + gauss(engine);
}

// Squared differences are formed in float and summed in double; peak 1.0.
double lpa_psnr(const float *a, const float *b, long long n)
{
    double sum = 0.0;
    for (long long i = 0; i < n; ++i) {
        const float d = a[i] - b[i];
        const float d2 = d * d;
        sum += (double)d2;
    }
    const double mse = sum / (double)n;
    return 10.0 * std::log10(1.0 / mse);
}

void lpa_synth_image(float *out, int width, int height, unsigned seed) { lpa::synth_image(out, width, height, seed); }

lpa_status lpa_load_image(const char *path, int *width, int *height, float *out, long long capacity)
{
    if (!path || !width || !height) return lpa::fail(LPA_EINVAL, "null argument");
    std::vector<float> img;
    int w = 0, h = 0;
    std::string err = lpa::load_png_gray(path, w, h, img);
    if (!err.empty()) return lpa::fail(err == "can't fopen" ? LPA_EIO : LPA_EINVAL, err + ": " + path);
    *width = w;
    *height = h;
    if (out) {
        if ((long long)img.size() > capacity) return lpa::fail(LPA_EINVAL, "output buffer too small");
        std::memcpy(out, img.data(), img.size() * sizeof(float));
    }
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_load_image_channels(const char *path, int channels, int *width, int *height, float *out,
                                   long long capacity)
{
    if (!path || !width || !height) return lpa::fail(LPA_EINVAL, "null argument");
    if (channels != 1 && channels != 3) return lpa::fail(LPA_EINVAL, "channels must be 1 or 3");
    std::vector<float> img;
    int w = 0, h = 0;
    std::string err = lpa::load_png(path, channels, w, h, img);
    if (!err.empty()) return lpa::fail(err == "can't fopen" ? LPA_EIO : LPA_EINVAL, err + ": " + path);
    *width = w;
    *height = h;
    if (out) {
        if ((long long)img.size() > capacity) return lpa::fail(LPA_EINVAL, "output buffer too small");
        std::memcpy(out, img.data(), img.size() * sizeof(float));
    }
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_write_hdr_channels(const char *path, int width, int height, int channels, const float *data)
{
    if (!path || !data || width <= 0 || height <= 0 || (channels != 1 && channels != 3))
        return lpa::fail(LPA_EINVAL, "bad argument");
    if (!lpa::write_hdr(path, width, height, channels, data))
        return lpa::fail(LPA_EIO, std::string("cannot write ") + path);
    lpa::clear_error();
    return LPA_OK;
}

lpa_status lpa_write_hdr(const char *path, int width, int height, const float *data)
{
    if (!path || !data || width <= 0 || height <= 0) return lpa::fail(LPA_EINVAL, "bad argument");
    if (!lpa::write_hdr(path, width, height, data)) return lpa::fail(LPA_EIO, std::string("cannot write ") + path);
    lpa::clear_error();
    return LPA_OK;
}

}  // extern "C"
