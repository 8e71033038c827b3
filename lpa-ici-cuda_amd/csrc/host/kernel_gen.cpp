// Kernel-set generation: directional LPA kernels of polynomial order m with
// uniform, Gaussian or square windows, and the writer of the reference's text
// format.
//
// The reference ships its sets precomputed (kernels/*.txt, generated with a
// modified LASIP demo_CreateLPAKernels.m that is not in the repository,
// readme.md:87-92, 239); the file names give the polynomial order
// (m_<m1>_<m2>), the direction count (d_<D>), the window lengths (h12_...)
// and the window (symmetric, gaussian, square or none).
//
// The construction here is the local polynomial approximation (LPA) of
// Katkovnik, Egiazarian and Astola: the zero-order estimator of a degree-m
// polynomial fitted by weighted least squares along the direction.
//   * Support (sector windows): the wedge of direction d of D and window
//     length h holds the lattice points (x, y) (x along image columns, y up,
//     i.e. y = -row) with (x, y) = (0, 0), or
//       |angle(x, y) - phi_d| <= pi / D  and  u = x cos(phi_d) + y sin(phi_d) < h - 1/2,
//     phi_d = 2 pi d / D, v = -x sin(phi_d) + y cos(phi_d).
//   * Fit: samples u = 0, 1, ..., h-1 along the direction with weights w(u);
//     P(u) = phi(u)^T Phi^{-1} e0, phi(u) = (1, u, ..., u^m'),
//     Phi = sum_u w(u) phi(u) phi(u)^T, m' = min(m, h - 1).
//   * Kernel: k(x, y) = w(u) w(v) P(u) on the wedge, divided by its sum.
//   * Windows: uniform (w = 1) or Gaussian, w(t) = exp(-t^2 / (2 (0.3 h)^2)).
//   * Square window (D = 4 only): the h x h quadrant u in [0, h-1],
//     v in [-(h-1), 0], uniform, fitted on all its points (each u column has
//     h points, so the fit is the same), no division needed.
// Weights are stored as the files store them, with six decimals ("%f").
//
// Checked against the shipped sets (tests/test_kernel_gen.py): every kernel of
// the square and symmetric (m = 0, uniform) sets, and every kernel of the
// axis directions (phi_d a multiple of 90 degrees) of the m = 1 and m = 2
// sets, uniform and Gaussian, equal the files weight for weight.  The
// oblique directions of the m >= 1 sets come out of the LASIP generator by a
// discretization the files do not pin down: the supports agree, the weights
// differ (for the Gaussian sets by at most 4.6 % of the kernel's largest
// weight).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lpa_ici_host.hpp"

namespace lpa {

namespace {

// Solve Phi c = e0 for a small symmetric positive definite Phi (n <= 5),
// Gaussian elimination with partial pivoting in double.
bool solve_e0(std::vector<double> a, int n, std::vector<double> &c)
{
    c.assign(n, 0.0);
    c[0] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(a[r * n + col]) > std::fabs(a[piv * n + col])) piv = r;
        if (std::fabs(a[piv * n + col]) < 1e-300) return false;
        if (piv != col) {
            for (int k = 0; k < n; ++k) std::swap(a[col * n + k], a[piv * n + k]);
            std::swap(c[col], c[piv]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double f = a[r * n + col] / a[col * n + col];
            for (int k = col; k < n; ++k) a[r * n + k] -= f * a[col * n + k];
            c[r] -= f * c[col];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double v = c[r];
        for (int k = r + 1; k < n; ++k) v -= a[r * n + k] * c[k];
        c[r] = v / a[r * n + r];
    }
    return true;
}

// P(u) of the weighted least-squares fit over samples (u_i, w_i): coefficients.
bool lpa_poly(const std::vector<double> &u, const std::vector<double> &w, int m, std::vector<double> &coef)
{
    const int n = m + 1;
    std::vector<double> phi((size_t)n * n, 0.0);
    for (size_t i = 0; i < u.size(); ++i) {
        double pi = 1.0;
        std::vector<double> p(n);
        for (int k = 0; k < n; ++k, pi *= u[i]) p[k] = pi;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) phi[a * n + b] += w[i] * p[a] * p[b];
    }
    return solve_e0(phi, n, coef);
}

double poly_at(const std::vector<double> &coef, double u)
{
    double v = 0.0;
    for (int k = (int)coef.size() - 1; k >= 0; --k) v = v * u + coef[k];
    return v;
}

// "%f" text of v, read back as the loader reads it.
float six_decimals(double v)
{
    char text[64];
    std::snprintf(text, sizeof text, "%f", v);
    return std::strtof(text, nullptr);
}

}  // namespace

std::string lpa_kernels(int D, int S, const int *h, int m, int window, std::vector<float> &w, std::vector<int> &off,
                        std::vector<int> &rows, std::vector<int> &cols)
{
    if (D < 1 || S < 1 || !h) return "kernel generation: need D >= 1 directions, S >= 1 scales and their lengths";
    if (window < 0 || window > 2) return "kernel generation: window must be 0 (uniform), 1 (Gaussian) or 2 (square)";
    if (m < 0 || m > 4) return "kernel generation: polynomial order must lie in 0..4";
    if (window == 2 && D != 4) return "kernel generation: the square window needs D = 4 (lattice-aligned quadrants)";
    if (window != 2 && D < 4)
        return "kernel generation: a wedge of half-width pi/D over 45 degrees does not fit its (2h-1)^2 matrix (D >= 4)";
    for (int s = 0; s < S; ++s)
        if (h[s] < 1 || h[s] > 32) return "kernel generation: window lengths must lie in 1..32 (kernels up to 63 x 63)";
    w.clear();
    off.clear();
    rows.clear();
    cols.clear();
    const double pi = std::acos(-1.0), half = pi / D;
    for (int d = 0; d < D; ++d) {
        const double phi = 2.0 * pi * d / D, c = std::cos(phi), sn = std::sin(phi);
        for (int s = 0; s < S; ++s) {
            const int r = h[s] - 1, n = 2 * r + 1;
            const double sigma = 0.3 * h[s];
            auto win = [&](double t) { return window == 1 ? std::exp(-t * t / (2.0 * sigma * sigma)) : 1.0; };
            // support and projections
            std::vector<int> idx;
            std::vector<double> us, vs;
            for (int row = -r; row <= r; ++row)
                for (int col = -r; col <= r; ++col) {
                    const double x = col, y = -row, u = x * c + y * sn, v = -x * sn + y * c;
                    bool take;
                    if (window == 2) {
                        take = u > -1e-9 && u < r + 1e-9 && v < 1e-9 && v > -r - 1e-9;
                    } else {
                        take = row == 0 && col == 0;
                        if (!take) {
                            const double a = std::remainder(std::atan2(y, x) - phi, 2.0 * pi);  // into [-pi, pi]
                            take = std::fabs(a) <= half + 1e-12 && u < h[s] - 0.5;
                        }
                    }
                    if (take) {
                        idx.push_back((row + r) * n + (col + r));
                        us.push_back(u);
                        vs.push_back(v);
                    }
                }
            const int mm = std::min(m, r);  // h samples fit at most degree h - 1
            std::vector<double> coef;
            if (window == 2) {
                if (!lpa_poly(us, std::vector<double>(us.size(), 1.0), mm, coef))
                    return "kernel generation: singular fit";
            } else {
                std::vector<double> su, sw;
                for (int k = 0; k <= r; ++k) {
                    su.push_back(k);
                    sw.push_back(win(k));
                }
                if (!lpa_poly(su, sw, mm, coef)) return "kernel generation: singular fit";
            }
            std::vector<double> val(idx.size());
            double sum = 0.0;
            for (size_t i = 0; i < idx.size(); ++i) {
                val[i] = win(us[i]) * win(vs[i]) * poly_at(coef, us[i]);
                sum += val[i];
            }
            if (window != 2)
                for (double &v : val) v /= sum;
            off.push_back((int)w.size());
            rows.push_back(n);
            cols.push_back(n);
            std::vector<float> dense((size_t)n * n, 0.0f);
            for (size_t i = 0; i < idx.size(); ++i) dense[idx[i]] = six_decimals(val[i]);
            w.insert(w.end(), dense.begin(), dense.end());
        }
    }
    return "";
}

std::string wedge_kernels(int D, int S, const int *h, std::vector<float> &w, std::vector<int> &off,
                          std::vector<int> &rows, std::vector<int> &cols)
{
    return lpa_kernels(D, S, h, 0, 0, w, off, rows, cols);
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
