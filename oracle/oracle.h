/*
 * oracle.h -- CPU restatement of AccRQA's RR / DET / LAM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP path and the "port" CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library
 * (accrqa_amd/libaccrqa_amd.so) never links or calls it.
 *
 * Every function restates one function of the reference CPU path
 * (src/AccRQA_CPU_function.cpp and src/AccRQA_functions.cpp of AccRQA 1.0.1)
 * with the same arithmetic order, so that recurrence decisions and line
 * length histograms are bit-identical to the reference.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - decision primitive: brute force against exact tie inputs and the IEEE
 *     closed form; golden fixtures from the compiled reference;
 *   - histograms: golden fixtures (raw h_D / h_V from the compiled reference),
 *     constant-series closed forms, run-length invariants, brute force;
 *   - measures: CRQA known-answer values held by the reference's tests,
 *     golden fixtures (hex-exact), closed forms;
 *   - Lmax search: brute-force "largest l with scan[l] > last" on random
 *     histograms.
 * No function is left parity unpinned.
 */
#ifndef ACCRQA_ORACLE_H
#define ACCRQA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* R(i,j) for one threshold (AccRQA_CPU_function.cpp:40-46). */
int oracle_R_element_f32(const float *x, int64_t i, int64_t j, float thr, int tau, int emb, int dist);
int oracle_R_element_f64(const double *x, int64_t i, int64_t j, double thr, int tau, int emb, int dist);

/* Full-matrix recurrence count (AccRQA_CPU_function.cpp:347-378, without the divide). */
int64_t oracle_rr_count_f32(const float *x, size_t n, float thr, int tau, int emb, int dist);
/* accrqa_RP (AccRQA_functions.cpp:996-1060): the N_c x N_c recurrence plot,
 * row-major, as bytes or ints; 0 or ERR_INVALID_ARGUMENT. */
int oracle_RP_u8_f32(uint8_t *out, const float *x, size_t n, float thr, int tau, int emb, int dist);
int oracle_RP_u8_f64(uint8_t *out, const double *x, size_t n, double thr, int tau, int emb, int dist);
int oracle_RP_i32_f32(int32_t *out, const float *x, size_t n, float thr, int tau, int emb, int dist);
int oracle_RP_i32_f64(int32_t *out, const double *x, size_t n, double thr, int tau, int emb, int dist);
int64_t oracle_rr_count_f64(const double *x, size_t n, double thr, int tau, int emb, int dist);

/* Diagonal line histogram h_D, size N_c+1, including the doubling and the
 * line-of-identity bin (AccRQA_CPU_function.cpp:587-642, 63-103, 156-162).
 * hist must hold N_c+1 zeros on entry. */
void oracle_det_hist_f32(uint64_t *hist, const float *x, size_t n, float thr, int tau, int emb, int dist);
void oracle_det_hist_f64(uint64_t *hist, const double *x, size_t n, double thr, int tau, int emb, int dist);

/* Horizontal (== vertical) line histogram h_V, size N_c+1
 * (AccRQA_CPU_function.cpp:446-490, 105-142).  hist must hold zeros. */
void oracle_lam_hist_f32(uint64_t *hist, const float *x, size_t n, float thr, int tau, int emb, int dist);
void oracle_lam_hist_f64(uint64_t *hist, const double *x, size_t n, double thr, int tau, int emb, int dist);

/* Suffix sums metric[k] = sum_{l>=k} l*h[l], scan[k] = sum_{l>=k} h[l]
 * (AccRQA_CPU_function.cpp:181-204). */
void oracle_process_hist(uint64_t *metric, uint64_t *scan, const uint64_t *hist, size_t hs);

/* Derived measures (AccRQA_functions.cpp:48-133). */
double oracle_primary(const uint64_t *metric, size_t lmin);
double oracle_secondary(const uint64_t *metric, const uint64_t *scan, size_t lmin);
double oracle_tertiary(const uint64_t *scan, size_t hs, int is_det);
double oracle_quaternary(const uint64_t *hist, const uint64_t *scan, size_t hs, size_t lmin);
double oracle_rr_value(const uint64_t *metric, int64_t nc);

/* Public-API restatements: same argument meaning, output packing and error
 * codes as accrqa_RR / accrqa_DET / accrqa_LAM on PLT_CPU
 * (AccRQA_functions.cpp:454-523, 710-778, 820-942).  Return the Accrqa_Error code. */
int oracle_RR_f32(float *out, const float *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                  const float *thr, int nthr, int dist);
int oracle_RR_f64(double *out, const double *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                  const double *thr, int nthr, int dist);
int oracle_DET_f32(float *out, const float *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                   const int *lmin, int nl, const float *thr, int nthr, int dist);
int oracle_DET_f64(double *out, const double *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                   const int *lmin, int nl, const double *thr, int nthr, int dist);
int oracle_LAM_f32(float *out, const float *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                   const int *vmin, int nv, const float *thr, int nthr, int dist);
int oracle_LAM_f64(double *out, const double *x, size_t n, const int *tau, int ntau, const int *emb, int nemb,
                   const int *vmin, int nv, const double *thr, int nthr, int dist);

/* Number of OpenMP threads the oracle's parallel loops use. */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
