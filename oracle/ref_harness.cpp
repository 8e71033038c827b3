// ORACLE — test infrastructure only.
//
// C entry points over the reference's own CPU program, compiled from its
// sources where they lie (src/v0_0_lpa_ici_2D_naive.cpp or
// src/v1_0_lpa_ici_2D_fused.cpp, with `main` renamed by -Dmain=ref_main) into
// oracle/_ref/libref_v0_0.so and oracle/_ref/libref_v1_0.so by oracle/Makefile.
// Nothing here re-implements the reference: each wrapper forwards to the
// reference function named beside it, so tests can compare our restatement,
// loader, noise and PSNR with the reference's own bits.
#include <cstring>
#include <string>
#include <vector>

// Reference prototypes (src/v0_0_lpa_ici_2D_naive.cpp:72-186, :230-239;
// src/v1_0_lpa_ici_2D_fused.cpp:72-186, :240-250).
bool load_LPA_kernels_standardized(const std::string &filename, std::vector<float> &host_kernel_weights,
                                   std::vector<int> &host_offsets, std::vector<int> &host_rows,
                                   std::vector<int> &host_cols, int &num_dirs, int &num_scales);
void compute_kernel_variances(std::vector<float> &host_kernel_weights, std::vector<int> &host_offsets,
                              std::vector<int> &host_rows, std::vector<int> &host_cols,
                              std::vector<float> &kernel_variances, float sigma_noise);
void create_noisy_image(float *img_clean, float *img_noisy, int img_size, float sigma_noise,
                        unsigned int noise_seed);
double calculate_PSNR(float *img_one, float *img_two, int img_size);
#ifdef REF_V1_0
void anisotropic_lpa_ici_cpu_fused(float *img_noisy, float *img_denoised, int width, int height, int channels,
                                   float *global_weights_buffer, int num_dirs, int num_scales,
                                   float *host_kernel_weights, int *host_offsets, int *host_rows,
                                   int *host_cols, float *kernel_variances, float ici_gamma);
#else
void anisotropic_lpa_ici_cpu_naive(float *img_noisy, float *img_denoised, int width, int height, int channels,
                                   float *global_weights_buffer, int num_dirs, int num_scales,
                                   float *host_kernel_weights, int *host_offsets, int *host_rows,
                                   int *host_cols, float *kernel_variances, float ici_gamma);
#endif
extern "C" float *stbi_loadf(char const *filename, int *x, int *y, int *comp, int req_comp);
extern "C" void stbi_image_free(void *retval_from_stbi_load);

namespace {
struct Loaded {
    std::vector<float> w;
    std::vector<int> off, rows, cols;
    int D = 0, S = 0;
};
Loaded g_loaded;
}  // namespace

extern "C" {

// load_LPA_kernels_standardized: returns 1 on success and keeps the arrays
// for ref_kernels_copy; 0 on failure (the reference prints the reason).
int ref_load_kernels(const char *path, int *D, int *S, int *n_weights)
{
    g_loaded = Loaded();
    if (!load_LPA_kernels_standardized(path, g_loaded.w, g_loaded.off, g_loaded.rows, g_loaded.cols,
                                       g_loaded.D, g_loaded.S))
        return 0;
    *D = g_loaded.D;
    *S = g_loaded.S;
    *n_weights = (int)g_loaded.w.size();
    return 1;
}

void ref_kernels_copy(float *w, int *off, int *rows, int *cols)
{
    std::memcpy(w, g_loaded.w.data(), g_loaded.w.size() * sizeof(float));
    std::memcpy(off, g_loaded.off.data(), g_loaded.off.size() * sizeof(int));
    std::memcpy(rows, g_loaded.rows.data(), g_loaded.rows.size() * sizeof(int));
    std::memcpy(cols, g_loaded.cols.data(), g_loaded.cols.size() * sizeof(int));
}

// compute_kernel_variances over the last loaded set.
void ref_variances(float sigma, float *out)
{
    std::vector<float> v(g_loaded.off.size(), 0.0f);
    compute_kernel_variances(g_loaded.w, g_loaded.off, g_loaded.rows, g_loaded.cols, v, sigma);
    std::memcpy(out, v.data(), v.size() * sizeof(float));
}

void ref_noisy(const float *clean, float *noisy, int n, float sigma, unsigned seed)
{
    std::vector<float> c(clean, clean + n);
    create_noisy_image(c.data(), noisy, n, sigma, seed);
}

double ref_psnr(const float *a, const float *b, int n)
{
    std::vector<float> x(a, a + n), y(b, b + n);
    return calculate_PSNR(x.data(), y.data(), n);
}

// The reference's L3 CPU orchestrator; out is zero-filled first as its
// caller does (src/v0_0_lpa_ici_2D_naive.cpp:370-371).
void ref_denoise(const float *noisy, float *out, int W, int H, int D, int S, const float *weights,
                 const int *offsets, const int *rows, const int *cols, const float *variances, float gamma)
{
    size_t N = (size_t)W * H;
    std::vector<float> x(noisy, noisy + N), wsum(N, 0.0f);
    int K = D * S;
    std::vector<int> off(offsets, offsets + K), r(rows, rows + K), c(cols, cols + K);
    size_t nw = 0;
    for (int k = 0; k < K; k++) nw = std::max(nw, (size_t)(offsets[k] + rows[k] * cols[k]));
    std::vector<float> w(weights, weights + nw), var(variances, variances + K);
    std::memset(out, 0, N * sizeof(float));
#ifdef REF_V1_0
    anisotropic_lpa_ici_cpu_fused(x.data(), out, W, H, 1, wsum.data(), D, S, w.data(), off.data(), r.data(),
                                  c.data(), var.data(), gamma);
#else
    anisotropic_lpa_ici_cpu_naive(x.data(), out, W, H, 1, wsum.data(), D, S, w.data(), off.data(), r.data(),
                                  c.data(), var.data(), gamma);
#endif
}

// stbi_loadf(path, .., req_comp=1), as the reference's main calls it.
int ref_loadf(const char *path, int *W, int *H, float *out, int cap)
{
    int w, h, c;
    float *p = stbi_loadf(path, &w, &h, &c, 1);
    if (!p) return 0;
    *W = w;
    *H = h;
    if (out && w * h <= cap) std::memcpy(out, p, sizeof(float) * (size_t)w * h);
    stbi_image_free(p);
    return 1;
}

}  // extern "C"
