/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference's LPA-ICI 2D denoise path, used as the
 * checker for the HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product never does.
 *
 * Build: gcc -O3 -ffp-contract=off (x86-64 baseline, no -march), so every
 * float operation is a separately rounded IEEE mul/add exactly as in the
 * reference's -O3 build (CMakeLists.txt:8), which has no FMA on baseline x86-64.
 *
 * Functions and what pins each one (tests/test_oracle.py):
 *   oracle_v0_0        src/v0_0_lpa_ici_2D_naive.cpp:785-814 (convolve2D) and
 *                      :817-929 (anisotropic_lpa_ici_cpu_naive).  Pinned
 *                      bit-exact against the compiled reference (oracle/_ref)
 *                      and the golden fixtures made from it.
 *   oracle_v1_0        src/v1_0_lpa_ici_2D_fused.cpp:803-846, 851-929.  Pinned
 *                      bit-exact against oracle_v0_0, the reference v1_0 build
 *                      and the fixtures.
 *   oracle_walk_f32    the early-exit scale walk over nonzero taps of
 *                      src/v4_1_lpa_ici_2D_sparse_shared.cu:1044-1091 in the
 *                      v0_0 summation order.  Pinned bit-exact against
 *                      oracle_v0_0 (adding a zero tap adds +-0, which changes
 *                      no value), and it also emits the scale maps, the
 *                      interval margins and per-direction estimates.
 *   oracle_walk_f64    the same walk in double with double variances (the
 *                      FP64 error reference).  Pinned by closed forms
 *                      (constant image, impulse, single direction) and by
 *                      agreement with the f32 walk to ~1e-6.
 *   oracle_walk_q      the fixed-point ring-sum walk, the HIP path's default
 *                      fast walk for kernel sets with one uniform weight per
 *                      kernel and nested supports (no reference counterpart:
 *                      it computes the reference's estimator w * sum(x) with
 *                      the sum in exact integers).  Pinned by the closed form
 *                      of a constant image, an impulse (orientation), the
 *                      f32 walk's tolerance bar and the FP64 walk.
 *   oracle_adjoint_f64 dL/dx and dL/dk of the frozen-selection forward (no
 *                      reference counterpart: the reference has no adjoint).
 *                      Pinned on its own, on the CPU, by central finite
 *                      differences of oracle_walk_f64 with frozen scales
 *                      (tests/test_oracle.py::test_adjoint_f64_central_differences):
 *                      dL/dx at 48 pixels (two-point differences, exact for a
 *                      function linear in x) and dL/dk at 48 dense taps (the
 *                      fourth-order difference; the output is rational in k
 *                      through sum w^2), on random frozen scale maps, for the
 *                      uniform symmetric set, the m=(2,0) gaussian set
 *                      (non-uniform, negative taps) and two square sets.
 *                      oracle_adjoint_f64_planted carries five plausible
 *                      mistakes the same test must catch (a wrong tap's weight
 *                      in the variance term, a rotated kernel in dL/dx, the
 *                      variance term's factor 2 dropped, y dropped from
 *                      yhat - y, border taps read as edge copies).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* v0_0: dense correlation of every scale over every pixel                   */
/* ------------------------------------------------------------------------ */

/* Correlation with zero padding, centre at (rows/2, cols/2), no kernel flip.
 * Follows src/v0_0_lpa_ici_2D_naive.cpp:785-814 loop for loop. */
static void convolve2d_dense(const float* img, float* out, int W, int H,
                             const float* k, int kr, int kc)
{
    for (int r = 0; r < H; r++) {
        for (int c = 0; c < W; c++) {
            float acc = 0.0f;
            for (int i = 0; i < kr; i++) {
                for (int j = 0; j < kc; j++) {
                    int ir = r - kr / 2 + i;
                    int ic = c - kc / 2 + j;
                    if (ir >= 0 && ir < H && ic >= 0 && ic < W)
                        acc += k[i * kc + j] * img[ir * W + ic];
                }
            }
            out[r * W + c] = acc;
        }
    }
}

static inline float fmaxf_ref(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline float fminf_ref(float a, float b) { return (b < a) ? b : a; } /* std::min */

/* Per-(pixel, direction) bookkeeping shared by the restatements: the selected
 * scale (uint8) and the smallest |U-L| over the decisions taken up to and
 * including the first break (the interval margin). */
static inline void note_margin(float* margin, size_t idx, float U, float L)
{
    if (!margin) return;
    double m = fabs((double)U - (double)L);
    if (m < (double)margin[idx]) margin[idx] = (float)m;
}

/* anisotropic_lpa_ici_cpu_naive, src/v0_0_lpa_ici_2D_naive.cpp:817-929.
 * out: W*H, overwritten (the reference's caller zero-fills it first).
 * scale_map / margin: optional [D][H][W]. */
int oracle_v0_0(const float* noisy, float* out, int W, int H, int D, int S,
                const float* weights, const int* offsets, const int* rows,
                const int* cols, const float* variances, float gamma,
                uint8_t* scale_map, float* margin)
{
    size_t N = (size_t)W * H;
    float* est_d = (float*)calloc(N, sizeof(float));
    float* var_d = (float*)calloc(N, sizeof(float));
    float* lo = (float*)malloc(N * sizeof(float));
    float* hi = (float*)malloc(N * sizeof(float));
    float* buf = (float*)calloc(N, sizeof(float));
    float* wsum = (float*)calloc(N, sizeof(float));
    uint8_t* broken = (uint8_t*)calloc(N, 1);
    if (!est_d || !var_d || !lo || !hi || !buf || !wsum || !broken) return -1;
    for (size_t i = 0; i < N; i++) out[i] = 0.0f;

    for (int d = 0; d < D; d++) {
        for (size_t i = 0; i < N; i++) {
            est_d[i] = 0.0f; var_d[i] = 0.0f; lo[i] = -INFINITY; hi[i] = INFINITY; broken[i] = 0;
            if (margin) margin[(size_t)d * N + i] = INFINITY;
        }
        for (int s = 0; s < S; s++) {
            int k = d * S + s;
            convolve2d_dense(noisy, buf, W, H, weights + offsets[k], rows[k], cols[k]);
            float var = variances[k];
            float thr = gamma * sqrtf(var);
            for (size_t i = 0; i < N; i++) {
                float lb = buf[i] - thr;
                float ub = buf[i] + thr;
                lo[i] = fmaxf_ref(lo[i], lb);
                hi[i] = fminf_ref(hi[i], ub);
                if (!broken[i]) note_margin(margin, (size_t)d * N + i, hi[i], lo[i]);
                if (hi[i] >= lo[i]) {
                    est_d[i] = buf[i];
                    var_d[i] = var;
                    if (scale_map) scale_map[(size_t)d * N + i] = (uint8_t)s;
                } else {
                    broken[i] = 1;
                }
            }
        }
        for (size_t i = 0; i < N; i++) {
            out[i] += est_d[i] * (1.0f / var_d[i]);
            wsum[i] += 1.0f / var_d[i];
        }
    }
    for (size_t i = 0; i < N; i++) out[i] = out[i] / wsum[i];

    free(est_d); free(var_d); free(lo); free(hi); free(buf); free(wsum); free(broken);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* v1_0: correlate only pixels whose interval is still non-empty             */
/* ------------------------------------------------------------------------ */

/* process_scale_fused + anisotropic_lpa_ici_cpu_fused,
 * src/v1_0_lpa_ici_2D_fused.cpp:803-846 and :851-929. */
int oracle_v1_0(const float* noisy, float* out, int W, int H, int D, int S,
                const float* weights, const int* offsets, const int* rows,
                const int* cols, const float* variances, float gamma)
{
    size_t N = (size_t)W * H;
    float* est_d = (float*)calloc(N, sizeof(float));
    float* var_d = (float*)calloc(N, sizeof(float));
    float* lo = (float*)malloc(N * sizeof(float));
    float* hi = (float*)malloc(N * sizeof(float));
    float* wsum = (float*)calloc(N, sizeof(float));
    if (!est_d || !var_d || !lo || !hi || !wsum) return -1;
    for (size_t i = 0; i < N; i++) out[i] = 0.0f;

    for (int d = 0; d < D; d++) {
        for (size_t i = 0; i < N; i++) {
            est_d[i] = 0.0f; var_d[i] = 0.0f; lo[i] = -INFINITY; hi[i] = INFINITY;
        }
        for (int s = 0; s < S; s++) {
            int k = d * S + s;
            const float* ker = weights + offsets[k];
            int kr = rows[k], kc = cols[k];
            float var = variances[k];
            float thr = gamma * sqrtf(var);
            for (int r = 0; r < H; r++) {
                for (int c = 0; c < W; c++) {
                    size_t i = (size_t)r * W + c;
                    if (!(hi[i] >= lo[i])) continue;
                    float acc = 0.0f;
                    for (int a = 0; a < kr; a++)
                        for (int b = 0; b < kc; b++) {
                            int ir = r - kr / 2 + a, ic = c - kc / 2 + b;
                            if (ir >= 0 && ir < H && ic >= 0 && ic < W)
                                acc += ker[a * kc + b] * noisy[ir * W + ic];
                        }
                    float lb = acc - thr, ub = acc + thr;
                    lo[i] = fmaxf_ref(lo[i], lb);
                    hi[i] = fminf_ref(hi[i], ub);
                    if (hi[i] >= lo[i]) { est_d[i] = acc; var_d[i] = var; }
                }
            }
        }
        for (size_t i = 0; i < N; i++) {
            out[i] += est_d[i] * (1.0f / var_d[i]);
            wsum[i] += 1.0f / var_d[i];
        }
    }
    for (size_t i = 0; i < N; i++) out[i] = out[i] / wsum[i];
    free(est_d); free(var_d); free(lo); free(hi); free(wsum);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Early-exit walk over the nonzero taps (fast oracle, same bits as v0_0)    */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;          /* nonzero taps */
    int* dr;        /* row offset from the kernel centre */
    int* dc;        /* column offset from the kernel centre */
    float* w;       /* weight */
} tap_list;

/* Nonzero taps of one dense kernel in row-major order (the order v0_0 adds
 * them in).  Zero taps are dropped: they add +-0 to the running sum. */
static int make_taps(const float* k, int kr, int kc, tap_list* t)
{
    int n = 0;
    for (int i = 0; i < kr * kc; i++) n += (k[i] != 0.0f);
    t->n = n;
    t->dr = (int*)malloc(sizeof(int) * (n ? n : 1));
    t->dc = (int*)malloc(sizeof(int) * (n ? n : 1));
    t->w = (float*)malloc(sizeof(float) * (n ? n : 1));
    if (!t->dr || !t->dc || !t->w) return -1;
    n = 0;
    for (int a = 0; a < kr; a++)
        for (int b = 0; b < kc; b++) {
            float v = k[a * kc + b];
            if (v != 0.0f) { t->dr[n] = a - kr / 2; t->dc[n] = b - kc / 2; t->w[n] = v; n++; }
        }
    return 0;
}

static void free_taps(tap_list* t, int n)
{
    for (int i = 0; i < n; i++) { free(t[i].dr); free(t[i].dc); free(t[i].w); }
    free(t);
}

static tap_list* make_all_taps(int K, const float* weights, const int* offsets,
                               const int* rows, const int* cols)
{
    tap_list* t = (tap_list*)calloc((size_t)K, sizeof(tap_list));
    if (!t) return NULL;
    for (int k = 0; k < K; k++)
        if (make_taps(weights + offsets[k], rows[k], cols[k], &t[k])) { free_taps(t, K); return NULL; }
    return t;
}

static inline float corr_f32(const float* img, int W, int H, int r, int c, const tap_list* t)
{
    float acc = 0.0f;
    for (int j = 0; j < t->n; j++) {
        int ir = r + t->dr[j], ic = c + t->dc[j];
        if (ir >= 0 && ir < H && ic >= 0 && ic < W) acc += t->w[j] * img[(size_t)ir * W + ic];
    }
    return acc;
}

/* Early-exit walk, bit-identical to oracle_v0_0 for y.
 * Optional outputs (each [D][H][W]): scale_map, margin, est (the selected
 * per-direction estimate).  Pixels are independent, so the row loop is split
 * over `threads` OpenMP threads without changing any bit. */
int oracle_walk_f32(const float* noisy, float* out, int W, int H, int D, int S,
                    const float* weights, const int* offsets, const int* rows,
                    const int* cols, const float* variances, float gamma,
                    uint8_t* scale_map, float* margin, float* est, int threads)
{
    int K = D * S;
    tap_list* taps = make_all_taps(K, weights, offsets, rows, cols);
    if (!taps) return -1;
    float* thr = (float*)malloc(sizeof(float) * K);
    float* inv = (float*)malloc(sizeof(float) * K);
    if (!thr || !inv) return -1;
    for (int k = 0; k < K; k++) { thr[k] = gamma * sqrtf(variances[k]); inv[k] = 1.0f / variances[k]; }
    size_t N = (size_t)W * H;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int r = 0; r < H; r++) {
        for (int c = 0; c < W; c++) {
            size_t i = (size_t)r * W + c;
            float num = 0.0f, den = 0.0f;
            for (int d = 0; d < D; d++) {
                float L = -INFINITY, U = INFINITY, best = 0.0f, bi = 0.0f;
                float m = INFINITY;
                int sel = 0;
                for (int s = 0; s < S; s++) {
                    int k = d * S + s;
                    float v = corr_f32(noisy, W, H, r, c, &taps[k]);
                    L = fmaxf_ref(L, v - thr[k]);
                    U = fminf_ref(U, v + thr[k]);
                    double mm = fabs((double)U - (double)L);
                    if (mm < (double)m) m = (float)mm;
                    if (U < L) break;
                    best = v; bi = inv[k]; sel = s;
                }
                num += best * bi;
                den += bi;
                if (scale_map) scale_map[(size_t)d * N + i] = (uint8_t)sel;
                if (margin) margin[(size_t)d * N + i] = m;
                if (est) est[(size_t)d * N + i] = best;
            }
            out[i] = num / den;
        }
    }
    free(thr); free(inv);
    free_taps(taps, K);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fixed-point ring-sum walk                                                 */
/* ------------------------------------------------------------------------ */

/* With q(v) = rintf(v * 2^F) and w_k the single nonzero weight of kernel k:
 *   yhat_k(p) = fl( fl((float) sum_{t in support k} q(x[p+t])) * (w_k * 2^-F) ),
 * the sum exact in integers (taps outside the image add q(0) = 0; the order
 * of the taps does not matter), then the interval update with lo and hi
 * rounded separately, the early exit and the aggregation of oracle_walk_f32.
 * Returns -2 if some |x| >= range or x is not finite (the fixed point does
 * not apply), -3 if a kernel's nonzero weights are not all equal, -1 on
 * allocation failure.  Optional outputs [D][H][W]: scale_map, est. */
int oracle_walk_q(const float* noisy, float* out, int W, int H, int D, int S,
                  const float* weights, const int* offsets, const int* rows,
                  const int* cols, const float* variances, float gamma,
                  int frac_bits, float range, uint8_t* scale_map, float* est, int threads)
{
    int K = D * S;
    size_t N = (size_t)W * H;
    for (size_t i = 0; i < N; i++)
        if (!(fabsf(noisy[i]) < range)) return -2;
    tap_list* taps = make_all_taps(K, weights, offsets, rows, cols);
    if (!taps) return -1;
    float* thr = (float*)malloc(sizeof(float) * K);
    float* inv = (float*)malloc(sizeof(float) * K);
    float* wq = (float*)malloc(sizeof(float) * K);
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * N);
    if (!thr || !inv || !wq || !q) return -1;
    const float qscale = ldexpf(1.0f, frac_bits), qinv = ldexpf(1.0f, -frac_bits);
    for (int k = 0; k < K; k++) {
        thr[k] = gamma * sqrtf(variances[k]);
        inv[k] = 1.0f / variances[k];
        for (int j = 0; j < taps[k].n; j++)
            if (memcmp(&taps[k].w[j], &taps[k].w[0], sizeof(float)) != 0) { free_taps(taps, K); return -3; }
        wq[k] = (taps[k].n ? taps[k].w[0] : 0.0f) * qinv;
    }
    for (size_t i = 0; i < N; i++) q[i] = (int32_t)rintf(noisy[i] * qscale);
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int r = 0; r < H; r++) {
        for (int c = 0; c < W; c++) {
            size_t i = (size_t)r * W + c;
            float num = 0.0f, den = 0.0f;
            for (int d = 0; d < D; d++) {
                float L = -INFINITY, U = INFINITY, best = 0.0f, bi = 0.0f;
                int sel = 0;
                for (int s = 0; s < S; s++) {
                    int k = d * S + s;
                    int64_t sum = 0;
                    for (int j = 0; j < taps[k].n; j++) {
                        int ir = r + taps[k].dr[j], ic = c + taps[k].dc[j];
                        if (ir >= 0 && ir < H && ic >= 0 && ic < W) sum += q[(size_t)ir * W + ic];
                    }
                    float v = (float)(int32_t)sum * wq[k];
                    L = fmaxf_ref(L, v - thr[k]);
                    U = fminf_ref(U, v + thr[k]);
                    if (U < L) break;
                    best = v; bi = inv[k]; sel = s;
                }
                num += best * bi;
                den += bi;
                if (scale_map) scale_map[(size_t)d * N + i] = (uint8_t)sel;
                if (est) est[(size_t)d * N + i] = best;
            }
            out[i] = num / den;
        }
    }
    free(thr); free(inv); free(wq); free(q);
    free_taps(taps, K);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 walk (error reference) and frozen-selection forward                  */
/* ------------------------------------------------------------------------ */

static inline double corr_f64(const float* img, int W, int H, int r, int c, const tap_list* t,
                              const double* wd)
{
    double acc = 0.0;
    for (int j = 0; j < t->n; j++) {
        int ir = r + t->dr[j], ic = c + t->dc[j];
        if (ir >= 0 && ir < H && ic >= 0 && ic < W) acc += wd[j] * (double)img[(size_t)ir * W + ic];
    }
    return acc;
}

/* Sum of squared weights of each dense kernel, in double. */
static void sumsq_f64(int K, const float* weights, const int* offsets, const int* rows,
                      const int* cols, double* out)
{
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        const float* w = weights + offsets[k];
        for (int i = 0; i < rows[k] * cols[k]; i++) s += (double)w[i] * (double)w[i];
        out[k] = s;
    }
}

/* The walk in double: var_k = sigma^2 * sum(w^2), thr = gamma*sqrt(var).
 * frozen (optional [D][H][W]): use these scales instead of walking; then the
 * output is the piecewise-linear map the adjoint differentiates.
 * scale_map (optional): the scales the double walk selects (or the frozen ones).
 * est (optional, double [D][H][W]): the selected per-direction estimates. */
int oracle_walk_f64(const float* noisy, double* out, int W, int H, int D, int S,
                    const float* weights, const int* offsets, const int* rows,
                    const int* cols, double sigma, double gamma,
                    const uint8_t* frozen, uint8_t* scale_map, double* est, int threads)
{
    int K = D * S;
    tap_list* taps = make_all_taps(K, weights, offsets, rows, cols);
    if (!taps) return -1;
    double* ss = (double*)malloc(sizeof(double) * K);
    double* var = (double*)malloc(sizeof(double) * K);
    double* thr = (double*)malloc(sizeof(double) * K);
    double** wd = (double**)malloc(sizeof(double*) * K);
    if (!ss || !var || !thr || !wd) return -1;
    sumsq_f64(K, weights, offsets, rows, cols, ss);
    for (int k = 0; k < K; k++) {
        var[k] = sigma * sigma * ss[k];
        thr[k] = gamma * sqrt(var[k]);
        wd[k] = (double*)malloc(sizeof(double) * (taps[k].n ? taps[k].n : 1));
        for (int j = 0; j < taps[k].n; j++) wd[k][j] = (double)taps[k].w[j];
    }
    size_t N = (size_t)W * H;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int r = 0; r < H; r++) {
        for (int c = 0; c < W; c++) {
            size_t i = (size_t)r * W + c;
            double num = 0.0, den = 0.0;
            for (int d = 0; d < D; d++) {
                double best = 0.0, bv = 1.0;
                int sel = 0;
                if (frozen) {
                    sel = frozen[(size_t)d * N + i];
                    int k = d * S + sel;
                    best = corr_f64(noisy, W, H, r, c, &taps[k], wd[k]);
                    bv = var[k];
                } else {
                    double L = -INFINITY, U = INFINITY;
                    for (int s = 0; s < S; s++) {
                        int k = d * S + s;
                        double v = corr_f64(noisy, W, H, r, c, &taps[k], wd[k]);
                        L = fmax(L, v - thr[k]);
                        U = fmin(U, v + thr[k]);
                        if (U < L) break;
                        best = v; bv = var[k]; sel = s;
                    }
                }
                num += best / bv;
                den += 1.0 / bv;
                if (scale_map) scale_map[(size_t)d * N + i] = (uint8_t)sel;
                if (est) est[(size_t)d * N + i] = best;
            }
            out[i] = num / den;
        }
    }
    for (int k = 0; k < K; k++) free(wd[k]);
    free(wd); free(ss); free(var); free(thr);
    free_taps(taps, K);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 adjoint of the frozen-selection forward                              */
/* ------------------------------------------------------------------------ */

/* With the scales s_d(p) frozen, let S_k = sum(w_k^2), w_d(p) = 1/(sigma^2 S),
 * Wt(p) = sum_d w_d(p), yhat_d(p) = sum_t k[t] x[p+t] and
 * y(p) = sum_d w_d yhat_d / Wt.  For an upstream gradient g:
 *   dL/dx[q]    = sum_p sum_d g(p) w_d(p)/Wt(p) * k_{d,s_d(p)}[q-p]
 *   dL/dk_{d,s}[t] = sum_{p: s_d(p)=s} g(p) w_d/Wt * (x[p+t] - 2 k[t] (yhat_d - y)/S)
 * Taps that fall outside the image read zero.  dk is laid out like the dense
 * weights (offsets/rows/cols) and is accumulated into (caller zero-fills).
 * dx is overwritten. */
static int adjoint_f64(const float* noisy, const float* dy, int W, int H, int D, int S,
                       const float* weights, const int* offsets, const int* rows,
                       const int* cols, double sigma, const uint8_t* scales,
                       double* dx, double* dk, int mistake)
{
    int K = D * S;
    double* ss = (double*)malloc(sizeof(double) * K);
    if (!ss) return -1;
    sumsq_f64(K, weights, offsets, rows, cols, ss);
    size_t N = (size_t)W * H;
    for (size_t i = 0; i < N; i++) dx[i] = 0.0;
    double* yhat = (double*)malloc(sizeof(double) * D);
    double* wdir = (double*)malloc(sizeof(double) * D);
    if (!yhat || !wdir) return -1;
    for (int r = 0; r < H; r++) {
        for (int c = 0; c < W; c++) {
            size_t i = (size_t)r * W + c;
            double num = 0.0, Wt = 0.0;
            for (int d = 0; d < D; d++) {
                int k = d * S + scales[(size_t)d * N + i];
                const float* w = weights + offsets[k];
                int kr = rows[k], kc = cols[k];
                double acc = 0.0;
                for (int a = 0; a < kr; a++)
                    for (int b = 0; b < kc; b++) {
                        int ir = r - kr / 2 + a, ic = c - kc / 2 + b;
                        if (ir >= 0 && ir < H && ic >= 0 && ic < W)
                            acc += (double)w[a * kc + b] * (double)noisy[(size_t)ir * W + ic];
                    }
                yhat[d] = acc;
                wdir[d] = 1.0 / (sigma * sigma * ss[k]);
                num += wdir[d] * acc;
                Wt += wdir[d];
            }
            double y = num / Wt;
            double g = (double)dy[i];
            for (int d = 0; d < D; d++) {
                int k = d * S + scales[(size_t)d * N + i];
                const float* w = weights + offsets[k];
                int kr = rows[k], kc = cols[k];
                double cd = g * wdir[d] / Wt;
                double corr = (mistake == 3 ? 1.0 : 2.0) * (yhat[d] - (mistake == 4 ? 0.0 : y)) / ss[k];
                double* gk = dk + offsets[k];
                for (int a = 0; a < kr; a++)
                    for (int b = 0; b < kc; b++) {
                        int ir = r - kr / 2 + a, ic = c - kc / 2 + b;
                        double xv = 0.0;
                        int inside = (ir >= 0 && ir < H && ic >= 0 && ic < W);
                        if (inside) {
                            xv = (double)noisy[(size_t)ir * W + ic];
                        } else if (mistake == 5) {
                            int er = ir < 0 ? 0 : ir >= H ? H - 1 : ir, ec = ic < 0 ? 0 : ic >= W ? W - 1 : ic;
                            xv = (double)noisy[(size_t)er * W + ec];
                        }
                        double kv = (double)w[a * kc + b];
                        double kvar = mistake == 1 ? (double)w[(a * kc + b + 1) % (kr * kc)] : kv;
                        double kdx = mistake == 2 ? (double)w[(kr - 1 - a) * kc + (kc - 1 - b)] : kv;
                        gk[a * kc + b] += cd * (xv - kvar * corr);
                        if (inside) dx[(size_t)ir * W + ic] += cd * kdx;
                    }
            }
        }
    }
    free(yhat); free(wdir); free(ss);
    return 0;
}

int oracle_adjoint_f64(const float* noisy, const float* dy, int W, int H, int D, int S,
                       const float* weights, const int* offsets, const int* rows,
                       const int* cols, double sigma, const uint8_t* scales,
                       double* dx, double* dk)
{
    return adjoint_f64(noisy, dy, W, H, D, S, weights, offsets, rows, cols, sigma, scales, dx, dk, 0);
}

/* The same with a planted mistake (1..5, see the header), for the test that
 * shows the finite-difference check catches each; 0 is the adjoint itself. */
int oracle_adjoint_f64_planted(const float* noisy, const float* dy, int W, int H, int D, int S,
                               const float* weights, const int* offsets, const int* rows,
                               const int* cols, double sigma, const uint8_t* scales,
                               double* dx, double* dk, int mistake)
{
    return adjoint_f64(noisy, dy, W, H, D, S, weights, offsets, rows, cols, sigma, scales, dx, dk, mistake);
}

/* Useful taps of the walk: sum over pixels and directions of the nonzero
 * taps of scales 0..min(s*+1, S-1) (the walk evaluates the first rejected
 * scale to see the break).  nnz: [D*S] nonzero taps per kernel. */
long long oracle_useful_taps(const uint8_t* scale_map, size_t N, int D, int S, const int* nnz)
{
    long long cum[256 * 16];
    long long total = 0;
    if (S > 16 || D > 256) return -1;
    for (int d = 0; d < D; d++) {
        long long acc = 0;
        for (int s = 0; s < S; s++) { acc += nnz[d * S + s]; cum[d * 16 + s] = acc; }
    }
    for (int d = 0; d < D; d++) {
        const uint8_t* m = scale_map + (size_t)d * N;
        for (size_t i = 0; i < N; i++) {
            int s = m[i] + 1;
            if (s > S - 1) s = S - 1;
            total += cum[d * 16 + s];
        }
    }
    return total;
}
