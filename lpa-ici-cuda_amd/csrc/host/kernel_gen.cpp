// Kernel-set generation: uniform directional wedges, and the writer of the
// reference's text format.
//
// The reference ships its sets precomputed (kernels/*.txt, generated with a
// modified LASIP demo_CreateLPAKernels.m that is not in the repository,
// readme.md:239).  Its default set, lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt
// (Config.hpp:9), is a family of uniform wedges: kernel (d, s) has the value
// 1/n on n lattice points and 0 elsewhere.  The wedge of direction d of D and
// window length h holds the lattice points (x, y) (x along image columns, y
// up, i.e. y = -row) with
//   (x, y) = (0, 0), or
//   |angle(x, y) - phi_d| <= pi / D   and   x cos(phi_d) + y sin(phi_d) < h - 1/2,
// phi_d = 2 pi d / D: a sector of half-width pi / D around the direction whose
// points project onto the direction's axis within h - 1/2 of the centre
// (the projection rounds to at most h - 1).  The kernel matrix is
// (2h - 1) x (2h - 1) with the centre tap in its middle (every wedge of
// half-width <= 45 degrees fits), and the weights are 1/n as the files store
// them, with six decimals ("%f").
//
// Generated this way, D = 16 and h = 1, 2, 4, 8, 16, 24, 32 give the shipped
// default set weight for weight after the loader's crop.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lpa_ici_host.hpp"

namespace lpa {

std::string wedge_kernels(int D, int S, const int *h, std::vector<float> &w, std::vector<int> &off,
                          std::vector<int> &rows, std::vector<int> &cols)
{
    if (D < 1 || S < 1 || !h) return "wedge kernels: need D >= 1 directions, S >= 1 scales and their lengths";
    if (D < 4) return "wedge kernels: a wedge of half-width pi/D over 45 degrees does not fit its (2h-1)^2 matrix (D >= 4)";
    for (int s = 0; s < S; ++s)
        if (h[s] < 1 || h[s] > 32) return "wedge kernels: window lengths must lie in 1..32 (kernels up to 63 x 63)";
    w.clear();
    off.clear();
    rows.clear();
    cols.clear();
    const double pi = std::acos(-1.0), half = pi / D;
    for (int d = 0; d < D; ++d) {
        const double phi = 2.0 * pi * d / D, c = std::cos(phi), sn = std::sin(phi);
        for (int s = 0; s < S; ++s) {
            const int r = h[s] - 1, n = 2 * r + 1;
            std::vector<char> in((size_t)n * n, 0);
            int count = 0;
            for (int row = -r; row <= r; ++row)
                for (int col = -r; col <= r; ++col) {
                    const double x = col, y = -row;
                    bool take = row == 0 && col == 0;
                    if (!take) {
                        double a = std::atan2(y, x) - phi;
                        a = std::remainder(a, 2.0 * pi);  // into [-pi, pi]
                        take = std::fabs(a) <= half + 1e-12 && x * c + y * sn < h[s] - 0.5;
                    }
                    if (take) {
                        in[(size_t)(row + r) * n + (col + r)] = 1;
                        ++count;
                    }
                }
            char text[32];
            std::snprintf(text, sizeof text, "%f", 1.0 / count);
            const float v = std::strtof(text, nullptr);
            off.push_back((int)w.size());
            rows.push_back(n);
            cols.push_back(n);
            for (size_t i = 0; i < in.size(); ++i) w.push_back(in[i] ? v : 0.0f);
        }
    }
    return "";
}

std::string write_kernels(const std::string &path, int D, int S, const float *w, const int *off, const int *rows,
                          const int *cols)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return "Failed to open kernel file for writing: " + path;
    bool ok = std::fprintf(f, "%d %d\n", D, S) > 0;
    for (int k = 0; k < D * S && ok; ++k) {
        ok = std::fprintf(f, "%d %d\n", rows[k], cols[k]) > 0;
        const float *kw = w + off[k];
        for (int i = 0; i < rows[k] && ok; ++i) {
            for (int j = 0; j < cols[k] && ok; ++j) ok = std::fprintf(f, j ? " %f" : "%f", kw[i * cols[k] + j]) > 0;
            ok = ok && std::fputc('\n', f) != EOF;
        }
    }
    ok = std::fclose(f) == 0 && ok;
    return ok ? "" : "Failed to write kernel file: " + path;
}

}  // namespace lpa
