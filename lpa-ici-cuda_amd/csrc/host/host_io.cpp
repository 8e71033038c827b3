// Host substrate: kernel-file loader, variance table, noise, PSNR, synthetic
// images and image I/O.  Plain C++ (no HIP), built with -ffp-contract=off so
// every float expression rounds the way the reference's x86-64 build does.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "lpa_ici.h"
#include "lpa_ici_host.hpp"

namespace lpa {

bool read_text_file(const std::string &path, std::string &out)
{
    out.clear();
    bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz) {
        gzFile f = gzopen(path.c_str(), "rb");
        if (!f) return false;
        char buf[1 << 16];
        int n;
        while ((n = gzread(f, buf, sizeof buf)) > 0) out.append(buf, (size_t)n);
        bool ok = n == 0;
        gzclose(f);
        return ok;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// Loader semantics of src/v0_0_lpa_ici_2D_naive.cpp:450-535: the stream
// extraction operator does the number parsing (so "1.000000" and friends
// round exactly as in the reference), every (d,s) kernel is cropped about its
// centre to direction 0's size at the same scale, and the messages are the
// reference's.
std::string load_kernels(const std::string &path, std::vector<float> &w, std::vector<int> &off,
                         std::vector<int> &rows, std::vector<int> &cols, int &D, int &S)
{
    w.clear();
    off.clear();
    rows.clear();
    cols.clear();
    std::string text;
    if (!read_text_file(path, text)) return "Failed to open " + path;
    std::istringstream in(text);
    D = S = 0;
    in >> D >> S;
    if (!in || D <= 0 || S <= 0) return "Error: invalid header in " + path;
    if ((long long)D * S > (1 << 20)) return "Error: invalid header in " + path;

    std::vector<int> tr((size_t)S), tc((size_t)S);
    std::vector<float> raw;
    int offset = 0;
    for (int d = 0; d < D; ++d) {
        for (int s = 0; s < S; ++s) {
            int fr = 0, fc = 0;
            in >> fr >> fc;
            const std::string where = std::to_string(d) + ", " + std::to_string(s) + " in " + path;
            if (!in || fr <= 0 || fc <= 0 || (long long)fr * fc > (1LL << 26))
                return "Error: invalid size of kernel " + where;
            raw.assign((size_t)fr * fc, 0.0f);
            for (size_t i = 0; i < raw.size(); ++i) in >> raw[i];
            if (!in) return "Error: missing weights of kernel " + where;
            if (d == 0) {
                tr[(size_t)s] = fr;
                tc[(size_t)s] = fc;
            }
            const int r0 = (fr - tr[(size_t)s]) / 2, c0 = (fc - tc[(size_t)s]) / 2;
            if (r0 < 0 || c0 < 0)
                return "Error: Matrix at direction " + std::to_string(d) + ", scale " + std::to_string(s) +
                       " is too small to crop to the target size!";
            off.push_back(offset);
            rows.push_back(tr[(size_t)s]);
            cols.push_back(tc[(size_t)s]);
            for (int r = 0; r < tr[(size_t)s]; ++r)
                for (int c = 0; c < tc[(size_t)s]; ++c) w.push_back(raw[(size_t)(r + r0) * fc + (c + c0)]);
            offset += tr[(size_t)s] * tc[(size_t)s];
        }
    }
    return "";
}

// ---------------------------------------------------------------------------
// PNG (8/16-bit gray, gray+alpha, RGB, RGBA, palette; not interlaced)
// ---------------------------------------------------------------------------

static uint32_t be32(const unsigned char *p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

static int paeth(int a, int b, int c)
{
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// 8-bit PNG to float with stbi_loadf(path, &w, &h, &c, channels) semantics
// for channels 1 or 3 (interleaved).  One channel: RGB folds to
// (77R + 150G + 29B) >> 8 (stb_image.h:1746-1749); three: gray replicates,
// palette entries expand.  16-bit samples keep their high byte, alpha is
// dropped, and every channel goes through powf(v / 255, 2.2)
// (stbi__ldr_to_hdr, stb_image.h:1858-1880).
std::string load_png(const std::string &path, int channels, int &width, int &height, std::vector<float> &out)
{
    if (channels != 1 && channels != 3) return "only 1 or 3 channels";
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return "can't fopen";
    std::vector<unsigned char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return "not a PNG file";
    size_t pos = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte;
    while (pos + 12 <= b.size()) {
        uint32_t len = be32(&b[pos]);
        if (pos + 12 + (size_t)len > b.size()) return "corrupt PNG";
        const unsigned char *type = &b[pos + 4], *data = &b[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            W = be32(data);
            H = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!ch || W == 0 || H == 0 || W > 65535 || H > 65535) return "unsupported PNG header";
    if (interlace != 0) return "interlaced PNG not supported";
    if (!(depth == 8 || (depth == 16 && ctype != 3))) return "PNG bit depth not supported";
    const int bytes = depth / 8;
    const size_t bpp = (size_t)ch * bytes, stride = (size_t)W * bpp;
    std::vector<unsigned char> raw(H * (stride + 1));
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size())
        return "corrupt PNG data";
    std::vector<unsigned char> img((size_t)H * stride), zero(stride, 0);
    for (uint32_t y = 0; y < H; ++y) {
        const unsigned char *src = &raw[y * (stride + 1) + 1];
        unsigned char *dst = &img[y * stride];
        const unsigned char *up = y ? &img[(y - 1) * stride] : zero.data();
        int ft = raw[y * (stride + 1)];
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= bpp ? dst[i - bpp] : 0, c = i >= bpp ? up[i - bpp] : 0, u = up[i], v;
            switch (ft) {
            case 0: v = src[i]; break;
            case 1: v = src[i] + a; break;
            case 2: v = src[i] + u; break;
            case 3: v = src[i] + ((a + u) >> 1); break;
            case 4: v = src[i] + paeth(a, u, c); break;
            default: return "bad PNG filter";
            }
            dst[i] = (unsigned char)v;
        }
    }
    width = (int)W;
    height = (int)H;
    out.resize((size_t)W * H * channels);
    auto to_float = [](int v) { return std::pow((float)v / 255.0f, 2.2f) * 1.0f; };  // pow in float, scale 1
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const unsigned char *p = &img[i * bpp];
        auto sample = [&](int k) { return (int)p[k * bytes]; };  // high byte of 16-bit samples
        int rgb[3];
        if (ctype == 3) {
            size_t idx = p[0];
            if (3 * idx + 2 >= plte.size()) return "bad PNG palette index";
            rgb[0] = plte[3 * idx];
            rgb[1] = plte[3 * idx + 1];
            rgb[2] = plte[3 * idx + 2];
        } else if (ch >= 3) {
            rgb[0] = sample(0);
            rgb[1] = sample(1);
            rgb[2] = sample(2);
        } else {
            rgb[0] = rgb[1] = rgb[2] = sample(0);
        }
        if (channels == 1) {
            const int g = (ctype == 3 || ch >= 3) ? (rgb[0] * 77 + rgb[1] * 150 + rgb[2] * 29) >> 8 : rgb[0];
            out[i] = to_float(g);
        } else {
            for (int c = 0; c < 3; ++c) out[3 * i + c] = to_float(rgb[c]);
        }
    }
    return "";
}

std::string load_png_gray(const std::string &path, int &width, int &height, std::vector<float> &out)
{
    return load_png(path, 1, width, height, out);
}

// Radiance RGBE with a shared exponent per pixel (stb_image_write.h:639-654):
// the largest component m sets the exponent, every component is written as
// (unsigned char)(v * frexp(m) * 256 / m); m below 1e-32 (negatives
// included) writes 0.  One channel is replicated to R, G and B.
bool write_hdr(const std::string &path, int width, int height, int channels, const float *data)
{
    if (channels != 1 && channels != 3) return false;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "#?RADIANCE\n# Written by lpa_ici\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", height, width);
    std::vector<unsigned char> line((size_t)width * 4);
    bool ok = true;
    for (int y = 0; y < height && ok; ++y) {
        for (int x = 0; x < width; ++x) {
            const float *v = &data[((size_t)y * width + x) * channels];
            const float r = v[0], g = channels == 3 ? v[1] : v[0], b = channels == 3 ? v[2] : v[0];
            const float m = std::max(r, std::max(g, b));
            unsigned char *q = &line[(size_t)x * 4];
            if (!(m >= 1e-32f)) {
                q[0] = q[1] = q[2] = q[3] = 0;
            } else {
                int e;
                const float n = std::frexp(m, &e) * 256.0f / m;
                q[0] = (unsigned char)(r * n);
                q[1] = (unsigned char)(g * n);
                q[2] = (unsigned char)(b * n);
                q[3] = (unsigned char)(e + 128);
            }
        }
        ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
    }
    return std::fclose(f) == 0 && ok;
}

bool write_hdr(const std::string &path, int width, int height, const float *data)
{
    return write_hdr(path, width, height, 1, data);
}

// ---------------------------------------------------------------------------
// Synthetic images
// ---------------------------------------------------------------------------

// Painter's-order scene, every parameter drawn from raw std::mt19937 words
// (never from a library distribution, so the image depends only on the seed):
// a linear ramp background, then disks, star-shaped polygons and disks of
// oriented sinusoidal stripes (period 3..12 px), each confined to its
// bounding box, finally clamped to [0,1].  Shape sizes are absolute (8..128
// px radius) and their number grows with the area, so large images keep the
// same local statistics as small ones.
void synth_image(float *out, int width, int height, unsigned seed)
{
    std::mt19937 rng(seed);
    auto U = [&]() { return (double)rng() * (1.0 / 4294967296.0); };
    const size_t N = (size_t)width * height;
    std::vector<double> img(N);
    const double base = 0.25 + 0.3 * U(), gx = (U() - 0.5) * 0.5, gy = (U() - 0.5) * 0.5;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            img[(size_t)y * width + x] = base + gx * x / width + gy * y / height;

    const long long n_shapes = 6 + (long long)N / 6000;
    const double two_pi = 6.283185307179586;
    for (long long k = 0; k < n_shapes; ++k) {
        const int kind = (int)(rng() % 3);
        const double cx = U() * width, cy = U() * height, r = 8.0 + 120.0 * U() * U();
        const double val = U();
        const int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(width - 1, (int)std::ceil(cx + r));
        const int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(height - 1, (int)std::ceil(cy + r));
        if (kind == 0) {  // disk of constant value
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    double dx = x - cx, dy = y - cy;
                    if (dx * dx + dy * dy <= r * r) img[(size_t)y * width + x] = val;
                }
        } else if (kind == 1) {  // star-shaped polygon, 3..8 vertices
            const int nv = 3 + (int)(rng() % 6);
            double px[8], py[8];
            const double a0 = U() * two_pi;
            for (int v = 0; v < nv; ++v) {
                double a = a0 + two_pi * (v + 0.8 * U()) / nv, rr = r * (0.45 + 0.55 * U());
                px[v] = cx + rr * std::cos(a);
                py[v] = cy + rr * std::sin(a);
            }
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    bool in = false;  // even-odd crossing test at the pixel centre
                    for (int i = 0, j = nv - 1; i < nv; j = i++) {
                        if ((py[i] > y) != (py[j] > y) &&
                            x < (px[j] - px[i]) * (y - py[i]) / (py[j] - py[i]) + px[i])
                            in = !in;
                    }
                    if (in) img[(size_t)y * width + x] = val;
                }
        } else {  // disk of oriented stripes
            const double period = 3.0 + 9.0 * U(), theta = U() * two_pi, phase = U() * two_pi;
            const double amp = 0.1 + 0.3 * U(), mean = 0.3 + 0.4 * val;
            const double c = std::cos(theta) * two_pi / period, s = std::sin(theta) * two_pi / period;
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    double dx = x - cx, dy = y - cy;
                    if (dx * dx + dy * dy <= r * r)
                        img[(size_t)y * width + x] = mean + amp * std::sin(c * x + s * y + phase);
                }
        }
    }
    for (size_t i = 0; i < N; ++i) out[i] = (float)std::min(1.0, std::max(0.0, img[i]));
}

}  // namespace lpa

// ---------------------------------------------------------------------------
// Reference-named functions
// ---------------------------------------------------------------------------

bool load_LPA_kernels_standardized(const std::string &filename, std::vector<float> &host_kernel_weights,
                                   std::vector<int> &host_offsets, std::vector<int> &host_rows,
                                   std::vector<int> &host_cols, int &num_dirs, int &num_scales)
{
    std::string err = lpa::load_kernels(filename, host_kernel_weights, host_offsets, host_rows, host_cols,
                                        num_dirs, num_scales);
    if (!err.empty()) {
        std::cerr << err << std::endl;
        return false;
    }
    return true;
}

void compute_kernel_variances(std::vector<float> &host_kernel_weights, std::vector<int> &host_offsets,
                              std::vector<int> &host_rows, std::vector<int> &host_cols,
                              std::vector<float> &kernel_variances, float sigma_noise)
{
    std::cout << "Precomputing kernel variances \n";
    kernel_variances.resize(host_offsets.size());
    lpa_kernel_variances((int)host_offsets.size(), host_kernel_weights.data(), host_offsets.data(),
                         host_rows.data(), host_cols.data(), sigma_noise, kernel_variances.data());
}

void create_noisy_image(float *img_clean, float *img_noisy, int img_size, float sigma_noise,
                        unsigned int noise_seed)
{
    lpa_make_noisy(img_clean, img_noisy, img_size, sigma_noise, noise_seed);
}

double calculate_PSNR(float *img_one, float *img_two, int img_size)
{
    return lpa_psnr(img_one, img_two, img_size);
}
