// Host-side mirror of the reference's driver-facing functions, with the
// reference's names, argument meaning and error behaviour, so a program
// written against src/v4_1_lpa_ici_2D_sparse_shared.cu's main() links
// against this library unchanged.  Device work goes through the C ABI in
// include/lpa_ici.h.
#ifndef LPA_ICI_HOST_HPP
#define LPA_ICI_HOST_HPP

#include <string>
#include <vector>

// src/v0_0_lpa_ici_2D_naive.cpp:450-535.  Prints the reference's message to
// std::cerr and returns false on a missing file, bad header, bad kernel size,
// missing weights or a kernel smaller than direction 0's at its scale.
bool load_LPA_kernels_standardized(const std::string &filename, std::vector<float> &host_kernel_weights,
                                   std::vector<int> &host_offsets, std::vector<int> &host_rows,
                                   std::vector<int> &host_cols, int &num_dirs, int &num_scales);

// src/v0_0_lpa_ici_2D_naive.cpp:675-701.
void compute_kernel_variances(std::vector<float> &host_kernel_weights, std::vector<int> &host_offsets,
                              std::vector<int> &host_rows, std::vector<int> &host_cols,
                              std::vector<float> &kernel_variances, float sigma_noise);

// src/v0_0_lpa_ici_2D_naive.cpp:743-760.
void create_noisy_image(float *img_clean, float *img_noisy, int img_size, float sigma_noise,
                        unsigned int noise_seed);

// src/v0_0_lpa_ici_2D_naive.cpp:762-783.
double calculate_PSNR(float *img_one, float *img_two, int img_size);

// The L3 orchestrator, src/v4_1_lpa_ici_2D_sparse_shared.cu:1122-1244:
// host buffers in and out, synchronous, one image, channels must be 1.
// host_global_weights_buffer is accepted and unused (the accumulators live in
// registers).  On any failure it prints the reason and exits with
// EXIT_FAILURE, as the reference does.
void anisotropic_lpa_ici_gpu_fused(float *host_img_noisy, float *host_img_denoised, int width, int height,
                                   int channels, float *host_global_weights_buffer, int num_dirs,
                                   int num_scales, float *host_kernel_weights, int host_kernel_weights_size,
                                   int *host_offsets, int *host_rows, int *host_cols,
                                   float *host_kernel_variances, float ici_gamma);

// Per-direction debugging outputs of the reference's v0_0 (commented out
// there, src/v0_0_lpa_ici_2D_naive.cpp:906-914): the selected estimate and
// the scale map s/(S-1) of every direction as .hdr files in `dir`.  Exits
// with EXIT_FAILURE on any failure, like the orchestrator.
void write_direction_outputs(const float *host_img_noisy, int width, int height, int num_dirs, int num_scales,
                             const float *host_kernel_weights, const int *host_offsets, const int *host_rows,
                             const int *host_cols, const float *host_kernel_variances, float ici_gamma,
                             const std::string &dir);

// Multichannel form of the orchestrator (the reference accepts one channel
// only): the interleaved channels are split into planes and denoised
// together in one persistent launch over the C ABI, each with the same
// kernel set, variances and gamma.  Prints the reference's timing line for
// the same window.  Exits with EXIT_FAILURE on any failure.
void lpa_ici_denoise_channels(const float *host_img_noisy, float *host_img_denoised, int width, int height,
                              int channels, int num_dirs, int num_scales, const float *host_kernel_weights,
                              const int *host_offsets, const int *host_rows, const int *host_cols,
                              const float *host_kernel_variances, float ici_gamma);

namespace lpa {

// Read a whole text file, transparently gunzipping names ending in ".gz".
bool read_text_file(const std::string &path, std::string &out);

// The loader's core: returns "" on success or the reference's message.
std::string load_kernels(const std::string &path, std::vector<float> &w, std::vector<int> &off,
                         std::vector<int> &rows, std::vector<int> &cols, int &D, int &S);

// 8-bit PNG to gray float (stbi_loadf(.., 1) semantics); "" or an error.
std::string load_png_gray(const std::string &path, int &width, int &height, std::vector<float> &out);

// 8-bit PNG to float with stbi_loadf(.., channels) semantics (1 or 3).
std::string load_png(const std::string &path, int channels, int &width, int &height, std::vector<float> &out);

// Radiance RGBE writer (shared exponent), 1 or 3 interleaved channels;
// pixels whose largest component is below 1e-32 (negatives included) are 0.
bool write_hdr(const std::string &path, int width, int height, const float *data);
bool write_hdr(const std::string &path, int width, int height, int channels, const float *data);

// Deterministic synthetic image in [0,1].
void synth_image(float *out, int width, int height, unsigned seed);

// Directional LPA kernel set (kernel_gen.cpp): D directions, S window
// lengths h[0..S), polynomial order m, window 0 uniform / 1 Gaussian / 2
// square; dense (2h-1)^2 kernels in the loader's layout.  "" or an error.
std::string lpa_kernels(int D, int S, const int *h, int m, int window, std::vector<float> &w, std::vector<int> &off,
                        std::vector<int> &rows, std::vector<int> &cols);

// Uniform wedge kernel set: lpa_kernels with m = 0 and the uniform window.
std::string wedge_kernels(int D, int S, const int *h, std::vector<float> &w, std::vector<int> &off,
                          std::vector<int> &rows, std::vector<int> &cols);

// The reference's kernel text format: "D S", then per kernel "rows cols" and
// its rows of "%f" weights.  "" or an error.
std::string write_kernels(const std::string &path, int D, int S, const float *w, const int *off, const int *rows,
                          const int *cols);

}  // namespace lpa

#endif
