// The reference's L3 orchestrator over the C ABI: host buffers in and out,
// synchronous, one image (src/v4_1_lpa_ici_2D_sparse_shared.cu:1122-1244).
// The reported time spans the same window as the reference's GpuTimer
// (src/Timer_gpu.cu:1-23): from after the image and the kernel set reach the
// device to after the result is back on the host.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "lpa_ici.h"
#include "lpa_ici_host.hpp"

namespace {

[[noreturn]] void die(const std::string &what)
{
    std::cerr << what << "\n";
    std::exit(EXIT_FAILURE);
}

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) die(std::string(hipGetErrorString(e)) + " in " + what);
}

void check(lpa_status st, const char *what)
{
    if (st != LPA_OK) die(std::string("Error in ") + what + ": " + lpa_last_error());
}

}  // namespace

void anisotropic_lpa_ici_gpu_fused(float *host_img_noisy, float *host_img_denoised, int width, int height,
                                   int channels, float * /*host_global_weights_buffer*/, int num_dirs,
                                   int num_scales, float *host_kernel_weights, int /*host_kernel_weights_size*/,
                                   int *host_offsets, int *host_rows, int *host_cols, float *host_kernel_variances,
                                   float ici_gamma)
{
    std::cout << "Executing LPA ICI fused algorithm \n";
    if (channels != 1) die("Error: only single-channel images are supported");
    const size_t bytes = sizeof(float) * (size_t)width * height;

    lpa_kset_t ks = nullptr;
    check(lpa_kset_create(num_dirs, num_scales, host_kernel_weights, host_offsets, host_rows, host_cols, &ks),
          "kernel set");
    float *d_x = nullptr, *d_y = nullptr;
    check(hipMalloc(&d_x, bytes), "hipMalloc");
    check(hipMalloc(&d_y, bytes), "hipMalloc");
    check(hipMemcpy(d_x, host_img_noisy, bytes, hipMemcpyHostToDevice), "hipMemcpy");

    lpa_image img{};
    img.x = d_x;
    img.y = d_y;
    img.width = width;
    img.height = height;
    img.gamma = ici_gamma;
    img.variances = host_kernel_variances;
    lpa_plan_t plan = nullptr;
    check(lpa_plan_create(ks, &img, 1, LPA_FWD_PLAIN, &plan), "plan");

    hipEvent_t t0, t1;
    check(hipEventCreate(&t0), "hipEventCreate");
    check(hipEventCreate(&t1), "hipEventCreate");
    check(hipEventRecord(t0, nullptr), "hipEventRecord");
    check(lpa_plan_run(plan, nullptr), "forward");
    check(hipMemcpy(host_img_denoised, d_y, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipEventRecord(t1, nullptr), "hipEventRecord");
    check(hipEventSynchronize(t1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    std::cout << "Time taken for directional loops + aggregation: " << ms / 1000.0f << " seconds\n";

    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    lpa_plan_destroy(plan);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    lpa_kset_destroy(ks);
}

// The reference's per-direction debugging outputs (commented out at
// src/v0_0_lpa_ici_2D_naive.cpp:906-914): for each direction d,
// <dir>gray_image_denoised_theta_<d>.hdr (the estimate at the selected scale)
// and <dir>scale_map_theta_<d>.hdr (s / (S - 1), as the reference's
// scale_map_d, :848 and :901), from one forward that keeps the uint8 scale
// map and each direction's selected estimate (the values it aggregated).
void write_direction_outputs(const float *host_img_noisy, int width, int height, int num_dirs, int num_scales,
                             const float *host_kernel_weights, const int *host_offsets, const int *host_rows,
                             const int *host_cols, const float *host_kernel_variances, float ici_gamma,
                             const std::string &dir)
{
    const size_t npx = (size_t)width * height;
    lpa_kset_t ks = nullptr;
    check(lpa_kset_create(num_dirs, num_scales, host_kernel_weights, host_offsets, host_rows, host_cols, &ks),
          "kernel set");
    float *d_x = nullptr, *d_y = nullptr, *d_est = nullptr;
    uint8_t *d_map = nullptr;
    check(hipMalloc(&d_x, sizeof(float) * npx), "hipMalloc");
    check(hipMalloc(&d_y, sizeof(float) * npx), "hipMalloc");
    check(hipMalloc(&d_est, sizeof(float) * npx * num_dirs), "hipMalloc");
    check(hipMalloc(&d_map, npx * num_dirs), "hipMalloc");
    check(hipMemcpy(d_x, host_img_noisy, sizeof(float) * npx, hipMemcpyHostToDevice), "hipMemcpy");
    lpa_image img{};
    img.x = d_x;
    img.y = d_y;
    img.scale_map = d_map;
    img.width = width;
    img.height = height;
    img.gamma = ici_gamma;
    img.variances = host_kernel_variances;
    img.estimates = d_est;  // the forward keeps the estimates it aggregates
    check(lpa_ici_forward(ks, &img, 1, LPA_FWD_PLAIN, nullptr), "forward");
    std::vector<float> h_est(npx * num_dirs), plane(npx);
    std::vector<uint8_t> h_map(npx * num_dirs);
    check(hipMemcpy(h_est.data(), d_est, sizeof(float) * h_est.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipMemcpy(h_map.data(), d_map, h_map.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    const std::string prefix = dir.empty() || dir.back() == '/' ? dir : dir + "/";
    for (int d = 0; d < num_dirs; ++d) {
        const std::string est_file = prefix + "gray_image_denoised_theta_" + std::to_string(d) + ".hdr";
        const std::string map_file = prefix + "scale_map_theta_" + std::to_string(d) + ".hdr";
        check(lpa_write_hdr(est_file.c_str(), width, height, h_est.data() + (size_t)d * npx), "write_hdr");
        for (size_t i = 0; i < npx; ++i)
            plane[i] = num_scales > 1 ? (float)h_map[(size_t)d * npx + i] / (num_scales - 1) : 0.0f;
        check(lpa_write_hdr(map_file.c_str(), width, height, plane.data()), "write_hdr");
    }
    std::cout << "Saved " << 2 * num_dirs << " per-direction estimates and scale maps in " << (dir.empty() ? "." : dir)
              << "\n";
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    (void)hipFree(d_est);
    (void)hipFree(d_map);
    lpa_kset_destroy(ks);
}

void lpa_ici_denoise_channels(const float *host_img_noisy, float *host_img_denoised, int width, int height,
                              int channels, int num_dirs, int num_scales, const float *host_kernel_weights,
                              const int *host_offsets, const int *host_rows, const int *host_cols,
                              const float *host_kernel_variances, float ici_gamma)
{
    std::cout << "Executing LPA ICI fused algorithm on " << channels << " channels\n";
    if (channels < 1) die("Error: channels must be positive");
    const size_t npx = (size_t)width * height;
    lpa_kset_t ks = nullptr;
    check(lpa_kset_create(num_dirs, num_scales, host_kernel_weights, host_offsets, host_rows, host_cols, &ks),
          "kernel set");
    std::vector<float> planes(npx * channels);
    for (size_t i = 0; i < npx; ++i)
        for (int c = 0; c < channels; ++c) planes[(size_t)c * npx + i] = host_img_noisy[i * channels + c];
    float *d_x = nullptr, *d_y = nullptr;
    check(hipMalloc(&d_x, sizeof(float) * planes.size()), "hipMalloc");
    check(hipMalloc(&d_y, sizeof(float) * planes.size()), "hipMalloc");
    check(hipMemcpy(d_x, planes.data(), sizeof(float) * planes.size(), hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<lpa_image> imgs(channels);
    for (int c = 0; c < channels; ++c) {
        lpa_image &img = imgs[c];
        img = lpa_image{};
        img.x = d_x + (size_t)c * npx;
        img.y = d_y + (size_t)c * npx;
        img.width = width;
        img.height = height;
        img.gamma = ici_gamma;
        img.variances = host_kernel_variances;
    }
    lpa_plan_t plan = nullptr;
    check(lpa_plan_create(ks, imgs.data(), channels, 0, &plan), "plan");
    hipEvent_t t0, t1;
    check(hipEventCreate(&t0), "hipEventCreate");
    check(hipEventCreate(&t1), "hipEventCreate");
    check(hipEventRecord(t0, nullptr), "hipEventRecord");
    check(lpa_plan_run(plan, nullptr), "forward");
    check(hipMemcpy(planes.data(), d_y, sizeof(float) * planes.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipEventRecord(t1, nullptr), "hipEventRecord");
    check(hipEventSynchronize(t1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    std::cout << "Time taken for directional loops + aggregation: " << ms / 1000.0f << " seconds\n";
    for (size_t i = 0; i < npx; ++i)
        for (int c = 0; c < channels; ++c) host_img_denoised[i * channels + c] = planes[(size_t)c * npx + i];
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    lpa_plan_destroy(plan);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    lpa_kset_destroy(ks);
}
