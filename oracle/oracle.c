/*
 * oracle.c -- CPU restatement of AccRQA's RR / DET / LAM path (C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the "port" CPU
 * baseline.  Never linked into the product library.
 *
 * Each function below names the reference function it restates
 * (src/AccRQA_CPU_function.cpp and src/AccRQA_functions.cpp, AccRQA 1.0.1).
 * The arithmetic order is the reference's: per component difference,
 * product and sum each rounded in the input type, no fused multiply-add
 * (this file is compiled with -ffp-contract=off), a correctly rounded square
 * root, and the decision (threshold - distance) >= 0 evaluated in the input
 * type.  The OpenMP structure (one pass per threshold, `omp for` over rows or
 * diagonals, thread-private histograms merged in a critical section) is the
 * reference's as well, so that timing this file is a fair stand-in for timing
 * the reference CPU path.
 *
 * Deliberate deviation (documented, also taken by the HIP path): inputs for
 * which the reference has undefined behaviour -- N_c = N - (m-1)*tau < 1,
 * tau < 0, m < 1, l_min outside [0, N_c] -- return ERR_INVALID_ARGUMENT (2)
 * instead of reading out of bounds.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ERR_SUCCESS 0
#define ERR_INVALID_ARGUMENT 2
#define ERR_MEM_ALLOC_FAILURE 4

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Per-type code, instantiated for float and double.
 *   distance_euclidean  AccRQA_CPU_function.cpp:14-24
 *   distance_maximum    AccRQA_CPU_function.cpp:27-37
 *   R_matrix_element    AccRQA_CPU_function.cpp:40-46
 * ---------------------------------------------------------------------- */
#define ORACLE_TYPED(F, SFX, SQRT, FABS)                                                          \
    static inline F dist_euclid_##SFX(const F *x, int64_t i, int64_t j, int tau, int emb) {       \
        F sum = 0;                                                                                \
        for (int k = 0; k < emb; k++) {                                                           \
            F a = x[i + (int64_t)k * tau];                                                        \
            F b = x[j + (int64_t)k * tau];                                                        \
            F d = a - b;                                                                          \
            F sq = d * d;                                                                         \
            sum = sum + sq;                                                                       \
        }                                                                                         \
        return SQRT(sum);                                                                         \
    }                                                                                             \
    static inline F dist_max_##SFX(const F *x, int64_t i, int64_t j, int tau, int emb) {          \
        F mx = 0;                                                                                 \
        for (int k = 0; k < emb; k++) {                                                           \
            F a = x[i + (int64_t)k * tau];                                                        \
            F b = x[j + (int64_t)k * tau];                                                        \
            F d = FABS(a - b);                                                                    \
            if (d > mx) mx = d;                                                                   \
        }                                                                                         \
        return mx;                                                                                \
    }                                                                                             \
    static inline int r_elem_##SFX(const F *x, int64_t i, int64_t j, F thr, int tau, int emb,     \
                                   int dist) {                                                    \
        F d = 0;                                                                                  \
        if (dist == 1) d = dist_euclid_##SFX(x, i, j, tau, emb);                                  \
        if (dist == 2) d = dist_max_##SFX(x, i, j, tau, emb);                                     \
        F diff = thr - d;                                                                         \
        return diff >= (F)0 ? 1 : 0;                                                              \
    }                                                                                             \
    int oracle_R_element_##SFX(const F *x, int64_t i, int64_t j, F thr, int tau, int emb,         \
                               int dist) {                                                        \
        return r_elem_##SFX(x, i, j, thr, tau, emb, dist);                                        \
    }                                                                                             \
    /* accrqa_RP, AccRQA_functions.cpp:996-1060 -> rqa_CPU_R_matrix_parallel,                      \
       AccRQA_CPU_function.cpp:226-244: out[i N_c + j] = R(i, j), row-major, char or int */        \
    int oracle_RP_u8_##SFX(uint8_t *out, const F *x, size_t n, F thr, int tau, int emb, int dist) {  \
        int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;                                       \
        if (nc < 1 || tau < 0 || emb < 1) return ERR_INVALID_ARGUMENT;                            \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t i = 0; i < nc; i++)                                                          \
            for (int64_t j = 0; j < nc; j++)                                                      \
                out[i * nc + j] = (uint8_t)r_elem_##SFX(x, i, j, thr, tau, emb, dist);            \
        return ERR_SUCCESS;                                                                       \
    }                                                                                             \
    int oracle_RP_i32_##SFX(int32_t *out, const F *x, size_t n, F thr, int tau, int emb, int dist) { \
        int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;                                       \
        if (nc < 1 || tau < 0 || emb < 1) return ERR_INVALID_ARGUMENT;                            \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t i = 0; i < nc; i++)                                                          \
            for (int64_t j = 0; j < nc; j++)                                                      \
                out[i * nc + j] = (int32_t)r_elem_##SFX(x, i, j, thr, tau, emb, dist);            \
        return ERR_SUCCESS;                                                                       \
    }                                                                                             \
    /* rqa_CPU_RR_metric_ref_parallel, AccRQA_CPU_function.cpp:347-378 (count only) */            \
    int64_t oracle_rr_count_##SFX(const F *x, size_t n, F thr, int tau, int emb, int dist) {      \
        int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;                                       \
        int64_t sum = 0;                                                                          \
        _Pragma("omp parallel for reduction(+:sum) schedule(static)")                             \
        for (int64_t i = 0; i < nc; i++) {                                                        \
            for (int64_t j = 0; j < nc; j++) sum += r_elem_##SFX(x, i, j, thr, tau, emb, dist);   \
        }                                                                                         \
        return sum;                                                                               \
    }                                                                                             \
    /* get_length_histogram_DET_inplace, AccRQA_CPU_function.cpp:63-103 */                        \
    static void det_diag_##SFX(uint64_t *h, const F *x, int64_t nc, int64_t r, F thr, int tau,    \
                               int emb, int dist) {                                               \
        int active = 0;                                                                           \
        int64_t len = 0;                                                                          \
        int64_t dl = nc - r;                                                                      \
        for (int64_t s = 0; s < dl; s++) {                                                        \
            int v = r_elem_##SFX(x, s, s + r, thr, tau, emb, dist);                               \
            if (v == 1 && active == 0) {                                                          \
                active = 1;                                                                       \
                len = 1;                                                                          \
            } else if (active == 1 && v == 0) {                                                   \
                h[len]++;                                                                         \
                active = 0;                                                                       \
                len = 0;                                                                          \
            } else if (active == 1 && v == 1) {                                                   \
                len++;                                                                            \
            }                                                                                     \
        }                                                                                         \
        if (active == 1) h[len]++;                                                                \
    }                                                                                             \
    /* rqa_CPU_DET_metric_parallel_mk1, AccRQA_CPU_function.cpp:587-642, and                      \
     * correctDETHistogram, AccRQA_CPU_function.cpp:156-162 */                                    \
    void oracle_det_hist_##SFX(uint64_t *hist, const F *x, size_t n, F thr, int tau, int emb,     \
                               int dist) {                                                        \
        int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;                                       \
        int64_t hs = nc + 1;                                                                      \
        _Pragma("omp parallel")                                                                   \
        {                                                                                         \
            uint64_t *loc = (uint64_t *)calloc((size_t)hs, sizeof(uint64_t));                     \
            _Pragma("omp for schedule(static) nowait")                                            \
            for (int64_t r = nc - 1; r > 0; r--) det_diag_##SFX(loc, x, nc, r, thr, tau, emb, dist); \
            _Pragma("omp critical")                                                               \
            {                                                                                     \
                for (int64_t f = 0; f < hs; f++) hist[f] += loc[f];                               \
            }                                                                                     \
            free(loc);                                                                            \
        }                                                                                         \
        for (int64_t f = 0; f < hs; f++) hist[f] = (uint64_t)(2.0 * (double)hist[f]);            \
        hist[hs - 1]++;                                                                           \
    }                                                                                             \
    /* get_length_histogram_LAM_inplace, AccRQA_CPU_function.cpp:105-142 */                       \
    static void lam_row_##SFX(uint64_t *h, const F *x, int64_t nc, int64_t row, F thr, int tau,   \
                              int emb, int dist) {                                                \
        int active = 0;                                                                           \
        int64_t len = 0;                                                                          \
        for (int64_t j = 0; j < nc; j++) {                                                        \
            int v = r_elem_##SFX(x, row, j, thr, tau, emb, dist);                                 \
            if (v == 1 && active == 0) {                                                          \
                active = 1;                                                                       \
                len = 1;                                                                          \
            } else if (active == 1 && v == 0) {                                                   \
                h[len]++;                                                                         \
                active = 0;                                                                       \
                len = 0;                                                                          \
            } else if (active == 1 && v == 1) {                                                   \
                len++;                                                                            \
            }                                                                                     \
        }                                                                                         \
        if (active == 1) h[len]++;                                                                \
    }                                                                                             \
    /* rqa_CPU_LAM_metric_parallel, AccRQA_CPU_function.cpp:446-490 */                            \
    void oracle_lam_hist_##SFX(uint64_t *hist, const F *x, size_t n, F thr, int tau, int emb,     \
                               int dist) {                                                        \
        int64_t nc = (int64_t)n - (int64_t)(emb - 1) * tau;                                       \
        int64_t hs = nc + 1;                                                                      \
        _Pragma("omp parallel")                                                                   \
        {                                                                                         \
            uint64_t *loc = (uint64_t *)calloc((size_t)hs, sizeof(uint64_t));                     \
            _Pragma("omp for schedule(static) nowait")                                            \
            for (int64_t i = 0; i < nc; i++) lam_row_##SFX(loc, x, nc, i, thr, tau, emb, dist);   \
            _Pragma("omp critical")                                                               \
            {                                                                                     \
                for (int64_t f = 0; f < hs; f++) hist[f] += loc[f];                               \
            }                                                                                     \
            free(loc);                                                                            \
        }                                                                                         \
    }

ORACLE_TYPED(float, f32, sqrtf, fabsf)
ORACLE_TYPED(double, f64, sqrt, fabs)

/* rqa_process_length_histogram with its helpers reverseArrayAndMultiply,
 * serialScanInclusive and reverseArray, AccRQA_CPU_function.cpp:145-204:
 * two suffix sums over the histogram. */
void oracle_process_hist(uint64_t *metric, uint64_t *scan, const uint64_t *hist, size_t hs) {
    uint64_t m = 0, s = 0;
    for (size_t k = hs; k-- > 0;) {
        m += (uint64_t)k * hist[k];
        s += hist[k];
        metric[k] = m;
        scan[k] = s;
    }
}

/* accrqaLengthHistogramResult::primary_metric, AccRQA_functions.cpp:48-57 */
double oracle_primary(const uint64_t *metric, size_t lmin) {
    if (metric[0] > 0) return (double)metric[lmin] / (double)metric[0];
    return NAN;
}

/* accrqaLengthHistogramResult::secondary_metric, AccRQA_functions.cpp:59-66 */
double oracle_secondary(const uint64_t *metric, const uint64_t *scan, size_t lmin) {
    if (scan[lmin] > 0) return (double)metric[lmin] / (double)scan[lmin];
    return NAN;
}

/* accrqaLengthHistogramResult::tertiary_metric, AccRQA_functions.cpp:68-112.
 * The reference's own bisection, step for step: it looks for the index mid
 * with scan[mid] == last and scan[mid-1] > last and returns mid-1. */
double oracle_tertiary(const uint64_t *scan, size_t hs, int is_det) {
    unsigned long long last, start, mid, end, length, mid_value, midm1_value;
    int found = 0;
    end = hs - 1;
    mid = hs / 2;
    last = is_det == 1 ? scan[end] : 0;
    start = 0;
    length = hs;
    while (!found) {
        mid_value = scan[mid];
        midm1_value = mid > 0 ? scan[mid - 1] : 0;
        if (mid_value == last && midm1_value > mid_value) {
            mid = mid - 1;
            found = 1;
        } else if (length == 0) {
            found = 1;
            break;
        } else if (mid == 0) {
            found = 1;
            break;
        } else if (mid_value > last) {
            start = mid;
            length = end - start;
            mid = mid + (length + 1) / 2;
        } else if (mid_value == last && midm1_value == last) {
            end = mid;
            length = end - start;
            mid = start + (length >> 1);
        }
    }
    return (double)mid;
}

/* accrqaLengthHistogramResult::quaternary_metric, AccRQA_functions.cpp:114-124 */
double oracle_quaternary(const uint64_t *hist, const uint64_t *scan, size_t hs, size_t lmin) {
    double entr = 0;
    for (size_t f = lmin; f < hs; f++) {
        double p = (double)hist[f] / (double)scan[lmin];
        if (p > 0) entr = entr + p * log(p);
    }
    return (-1.0) * entr;
}

/* accrqaLengthHistogramResult::RR_value, AccRQA_functions.cpp:126-133 */
double oracle_rr_value(const uint64_t *metric, int64_t nc) {
    return (double)metric[0] / (double)(nc * nc);
}

/* ------------------------------------------------------------------------
 * Public-API restatements.
 *   accrqa_LAM_CPU_t / calculate_LAM_CPU_default  AccRQA_functions.cpp:454-481, 513-523
 *   accrqa_DET_CPU_t / calculate_DET_CPU_default  AccRQA_functions.cpp:710-737, 768-778
 *   accrqa_RR_CPU_t  / calculate_RR_CPU_default   AccRQA_functions.cpp:820-856, 921-942
 * Output order: LAM threshold fastest (470); DET l_min fastest (726); RR (851).
 * ---------------------------------------------------------------------- */
static int check_common(size_t n, const int *tau, int ntau, const int *emb, int nemb, const int *lm, int nl,
                        int nthr) {
    for (int a = 0; a < ntau; a++) {
        if (tau[a] < 0) return ERR_INVALID_ARGUMENT;
        for (int b = 0; b < nemb; b++) {
            if (emb[b] < 1) return ERR_INVALID_ARGUMENT;
            int64_t nc = (int64_t)n - (int64_t)(emb[b] - 1) * tau[a];
            if (nc < 1) return ERR_INVALID_ARGUMENT;
            for (int c = 0; c < nl; c++)
                if (lm[c] < 0 || (int64_t)lm[c] > nc) return ERR_INVALID_ARGUMENT;
        }
    }
    (void)nthr;
    return ERR_SUCCESS;
}

#define ORACLE_API(F, SFX)                                                                        \
    int oracle_RR_##SFX(F *out, const F *x, size_t n, const int *tau, int ntau, const int *emb,   \
                        int nemb, const F *thr, int nthr, int dist) {                             \
        int err = ERR_SUCCESS;                                                                    \
        if (n == 0 || ntau <= 0 || nemb <= 0 || nthr <= 0) err = ERR_INVALID_ARGUMENT;            \
        if (!out || !x || !tau || !emb || !thr) err = ERR_MEM_ALLOC_FAILURE;                      \
        if (err) return err;                                                                      \
        if ((err = check_common(n, tau, ntau, emb, nemb, NULL, 0, nthr))) return err;             \
        for (int a = 0; a < ntau; a++)                                                            \
            for (int b = 0; b < nemb; b++)                                                        \
                for (int t = 0; t < nthr; t++) {                                                  \
                    int64_t nc = (int64_t)n - (int64_t)(emb[b] - 1) * tau[a];                     \
                    int64_t cnt = oracle_rr_count_##SFX(x, n, thr[t], tau[a], emb[b], dist);      \
                    int64_t tmp = nc * nc;                                                        \
                    out[a * nemb * nthr + b * nthr + t] = ((F)cnt) / ((F)tmp);                    \
                }                                                                                 \
        return ERR_SUCCESS;                                                                       \
    }                                                                                             \
    int oracle_DET_##SFX(F *out, const F *x, size_t n, const int *tau, int ntau, const int *emb,  \
                         int nemb, const int *lmin, int nl, const F *thr, int nthr, int dist) {   \
        int err = ERR_SUCCESS;                                                                    \
        if (n == 0 || ntau <= 0 || nemb <= 0 || nl <= 0 || nthr <= 0) err = ERR_INVALID_ARGUMENT; \
        if (!out || !x || !tau || !emb || !lmin || !thr) err = ERR_MEM_ALLOC_FAILURE;             \
        if (err) return err;                                                                      \
        if ((err = check_common(n, tau, ntau, emb, nemb, lmin, nl, nthr))) return err;            \
        for (int a = 0; a < ntau; a++)                                                            \
            for (int b = 0; b < nemb; b++)                                                        \
                for (int t = 0; t < nthr; t++) {                                                  \
                    int64_t nc = (int64_t)n - (int64_t)(emb[b] - 1) * tau[a];                     \
                    size_t hs = (size_t)nc + 1;                                                   \
                    uint64_t *h = (uint64_t *)calloc(hs, 8), *me = (uint64_t *)calloc(hs, 8);     \
                    uint64_t *sc = (uint64_t *)calloc(hs, 8);                                     \
                    oracle_det_hist_##SFX(h, x, n, thr[t], tau[a], emb[b], dist);                 \
                    oracle_process_hist(me, sc, h, hs);                                           \
                    for (int l = 0; l < nl; l++) {                                                \
                        size_t pos = (size_t)a * nl * nemb * nthr + (size_t)b * nl * nthr +       \
                                     (size_t)t * nl + l;                                          \
                        out[5 * pos + 0] = (F)oracle_primary(me, lmin[l]);                        \
                        out[5 * pos + 1] = (F)oracle_secondary(me, sc, lmin[l]);                  \
                        out[5 * pos + 2] = (F)oracle_tertiary(sc, hs, 1);                         \
                        out[5 * pos + 3] = (F)oracle_quaternary(h, sc, hs, lmin[l]);              \
                        out[5 * pos + 4] = (F)oracle_rr_value(me, nc);                            \
                    }                                                                             \
                    free(h);                                                                      \
                    free(me);                                                                     \
                    free(sc);                                                                     \
                }                                                                                 \
        return ERR_SUCCESS;                                                                       \
    }                                                                                             \
    int oracle_LAM_##SFX(F *out, const F *x, size_t n, const int *tau, int ntau, const int *emb,  \
                         int nemb, const int *vmin, int nv, const F *thr, int nthr, int dist) {   \
        int err = ERR_SUCCESS;                                                                    \
        if (n == 0 || ntau <= 0 || nemb <= 0 || nv <= 0 || nthr <= 0) err = ERR_INVALID_ARGUMENT; \
        if (!out || !x || !tau || !emb || !vmin || !thr) err = ERR_MEM_ALLOC_FAILURE;             \
        if (err) return err;                                                                      \
        if ((err = check_common(n, tau, ntau, emb, nemb, vmin, nv, nthr))) return err;            \
        for (int a = 0; a < ntau; a++)                                                            \
            for (int b = 0; b < nemb; b++)                                                        \
                for (int t = 0; t < nthr; t++) {                                                  \
                    int64_t nc = (int64_t)n - (int64_t)(emb[b] - 1) * tau[a];                     \
                    size_t hs = (size_t)nc + 1;                                                   \
                    uint64_t *h = (uint64_t *)calloc(hs, 8), *me = (uint64_t *)calloc(hs, 8);     \
                    uint64_t *sc = (uint64_t *)calloc(hs, 8);                                     \
                    oracle_lam_hist_##SFX(h, x, n, thr[t], tau[a], emb[b], dist);                 \
                    oracle_process_hist(me, sc, h, hs);                                           \
                    for (int v = 0; v < nv; v++) {                                                \
                        size_t pos = (size_t)a * nv * nemb * nthr + (size_t)b * nv * nthr +       \
                                     (size_t)v * nthr + t;                                        \
                        out[5 * pos + 0] = (F)oracle_primary(me, vmin[v]);                        \
                        out[5 * pos + 1] = (F)oracle_secondary(me, sc, vmin[v]);                  \
                        out[5 * pos + 2] = (F)oracle_tertiary(sc, hs, 0);                         \
                        out[5 * pos + 3] = (F)oracle_quaternary(h, sc, hs, vmin[v]);              \
                        out[5 * pos + 4] = (F)oracle_rr_value(me, nc);                            \
                    }                                                                             \
                    free(h);                                                                      \
                    free(me);                                                                     \
                    free(sc);                                                                     \
                }                                                                                 \
        return ERR_SUCCESS;                                                                       \
    }

ORACLE_API(float, f32)
ORACLE_API(double, f64)
