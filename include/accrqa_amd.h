/*
 * accrqa_amd.h -- C ABI of the MI355X (gfx950) RR / DET / LAM engine.
 *
 * This is the drop-in boundary: the reference's platform dispatch
 * (src/AccRQA_functions.cpp:525-555, 780-810, 945-975) gains one branch for
 * the new platform value ACCRQA_PLT_AMD_GPU that calls these functions in
 * place of its CUDA wrappers (src/AccRQA_GPU_function.hpp:14-30).  Plain C
 * types only; every buffer argument below is HOST memory unless its name
 * starts with d_ (device memory on the current HIP device).
 *
 * Error codes are the reference's Accrqa_Error values
 * (include/AccRQA_utilities_error.hpp:11-48): 0 success, 2 invalid argument,
 * 4 NULL array, 8 no GPU device, 9 not enough device memory, 10 other device
 * error.  As in the reference, *error is set to 0 on entry.
 *
 * Output layouts are the reference PLT_CPU layouts:
 *   RR : [nTaus][nEmbs][nThresholds]                      (AccRQA_functions.cpp:851)
 *   DET: [nTaus][nEmbs][nThresholds][nLmins][5]           (l_min fastest, :726)
 *   LAM: [nTaus][nEmbs][nVmins][nThresholds][5]           (threshold fastest, :470)
 * DET tuple = DET, L, Lmax, ENTR, RR; LAM tuple = LAM, TT, TTmax, ENTR, RR.
 * All five values are always computed (the reference CPU path ignores
 * calc_ENTR as well).
 *
 * Deviation from the reference (which has undefined behaviour there): a
 * (tau, emb) pair with N_c = N - (emb-1)*tau < 1, tau < 0, emb < 1, or an
 * l_min / v_min outside [0, N_c] returns 2 instead of reading out of bounds.
 * distance_type must be 1 (Euclidean) or 2 (maximal), as in the reference's
 * GPU wrappers.
 */
#ifndef ACCRQA_AMD_H
#define ACCRQA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* New value of Accrqa_CompPlatform (include/AccRQA_utilities_comp_platform.hpp:12-18
 * holds 0, 1 and 1024). */
#define ACCRQA_PLT_AMD_GPU 2048

/* Replaces the GPU branch of accrqa_RR(float*) (src/AccRQA_functions.cpp:945-959),
 * i.e. calculate_RR_GPU_default -> GPU_RQA_RR_metric_integer
 * (src/AccRQA_functions.cpp:858-890, src/AccRQA-GPU-RR.cu:399-414).
 * RR = (float)count / (float)(N_c*N_c), the PLT_CPU formula (AccRQA_CPU_function.cpp:376-377). */
void accrqa_amd_RR_f32(float *output, const float *input_data, size_t data_size, const int *tau_values,
                       int nTaus, const int *emb_values, int nEmbs, const float *threshold_values,
                       int nThresholds, int distance_type, int *error);
/* Replaces the GPU branch of accrqa_RR(double*) (src/AccRQA_functions.cpp:961-975). */
void accrqa_amd_RR_f64(double *output, const double *input_data, size_t data_size, const int *tau_values,
                       int nTaus, const int *emb_values, int nEmbs, const double *threshold_values,
                       int nThresholds, int distance_type, int *error);

/* Replaces the GPU branch of accrqa_DET(float*) (src/AccRQA_functions.cpp:780-794), i.e.
 * calculate_DET_GPU_default / calculate_DET_GPU_boxcar_square -> GPU_RQA_length_histogram_diagonal
 * (src/AccRQA_functions.cpp:568-706, src/AccRQA-GPU-HST.cu:2811-2817). */
void accrqa_amd_DET_f32(float *output, const float *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nLmins,
                        const float *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error);
/* Replaces the GPU branch of accrqa_DET(double*) (src/AccRQA_functions.cpp:796-810). */
void accrqa_amd_DET_f64(double *output, const double *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nLmins,
                        const double *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error);

/* Replaces the GPU branch of accrqa_LAM(float*) (src/AccRQA_functions.cpp:525-539), i.e.
 * calculate_LAM_GPU_default / calculate_LAM_GPU_boxcar_square -> GPU_RQA_length_histogram_vertical
 * (src/AccRQA_functions.cpp:391-452, src/AccRQA-GPU-HST.cu:2791-2867). */
void accrqa_amd_LAM_f32(float *output, const float *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *vmin_values, int nVmins,
                        const float *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error);
/* Replaces the GPU branch of accrqa_LAM(double*) (src/AccRQA_functions.cpp:541-555). */
void accrqa_amd_LAM_f64(double *output, const double *input_data, size_t data_size, const int *tau_values,
                        int nTaus, const int *emb_values, int nEmbs, const int *vmin_values, int nVmins,
                        const double *threshold_values, int nThresholds, int distance_type, int calc_ENTR,
                        int *error);

/* Raw integer results for one (tau, emb) and nThresholds thresholds.  Replaces the
 * histogram exports of the reference's GPU wrappers GPU_RQA_RR_metric_integer,
 * GPU_RQA_length_histogram_diagonal and GPU_RQA_length_histogram_vertical
 * (src/AccRQA_GPU_function.hpp:14-30), all thresholds in one call.
 *   which: bit 0 = rr_count[T] (full-matrix recurrence count),
 *          bit 1 = hist_det[T][N_c+1] (h_D, doubled, with the line-of-identity bin),
 *          bit 2 = hist_lam[T][N_c+1] (h_V).
 * Arrays for bits not set may be NULL. */
void accrqa_amd_histograms_f32(const float *input_data, size_t data_size, int tau, int emb,
                               const float *threshold_values, int nThresholds, int distance_type, int which,
                               unsigned long long *rr_count, unsigned long long *hist_det,
                               unsigned long long *hist_lam, int *error);
void accrqa_amd_histograms_f64(const double *input_data, size_t data_size, int tau, int emb,
                               const double *threshold_values, int nThresholds, int distance_type, int which,
                               unsigned long long *rr_count, unsigned long long *hist_det,
                               unsigned long long *hist_lam, int *error);

/* Device-resident form of one (tau, emb) block of the three measures, for callers
 * whose series already lives in HBM (no host<->device copy of the input).
 * d_input: N samples on the current device.  thresholds and lmins are host arrays.
 * Outputs are device arrays in the reference layout of one (tau, emb) block:
 *   d_rr [T], d_det [T][nL][5], d_lam [nL][T][5]   (any may be NULL to skip it).
 * stream: a hipStream_t (NULL = default stream).  Returns the error code.
 * The call is asynchronous with respect to the host; synchronise the stream
 * before reading the outputs.  Calls share the library's work buffers: the
 * library orders every call after the previous one on the device (the
 * stream of a call waits for the previous call's completion event when the
 * streams differ), so calls on several streams serialise rather than race. */
int accrqa_amd_rqa_device_f32(const float *d_input, size_t data_size, int tau, int emb,
                              const float *threshold_values, int nThresholds, const int *lmin_values, int nL,
                              int distance_type, float *d_rr, float *d_det, float *d_lam, void *stream);
int accrqa_amd_rqa_device_f64(const double *d_input, size_t data_size, int tau, int emb,
                              const double *threshold_values, int nThresholds, const int *lmin_values, int nL,
                              int distance_type, double *d_rr, double *d_det, double *d_lam, void *stream);

/* No counterpart in the reference.  accrqa_amd_rqa_device_f32 for callers
 * that cannot allocate device memory (code under HIP stream capture, tight
 * serving loops, harnesses that forbid hipMalloc): same arguments, outputs
 * and results, but the call allocates nothing and copies nothing from the
 * host.  Its work memory is a fixed array in device memory (8.2 MiB per GPU,
 * held until the process ends); the series must fit it: data_size <= 32768
 * samples, nThresholds <= 32, nL <= 16 (else 2, ERR_INVALID_ARGUMENT, with
 * nothing launched).  Every call on one GPU uses that GPU's array, so calls
 * that overlap on one GPU (several streams, several host threads) race; the
 * caller orders them. */
int accrqa_amd_rqa_device_static_f32(const float *d_input, size_t data_size, int tau, int emb,
                                     const float *threshold_values, int nThresholds, const int *lmin_values, int nL,
                                     int distance_type, float *d_rr, float *d_det, float *d_lam, void *stream);

/* Many series in one call (no counterpart in the reference, whose API takes
 * one series per call: the batched configuration's form).  Series k is
 * series[offsets[k] .. offsets[k+1]) (offsets has nSeries + 1 entries,
 * increasing); threshold_values is [nSeries][nThresholds] (thresholds per
 * series).  Outputs are host arrays in the per-series layouts of
 * accrqa_amd_RR / DET / LAM stacked over the series:
 *   rr_out  [nSeries][nTaus][nEmbs][nThresholds]
 *   det_out [nSeries][nTaus][nEmbs][nThresholds][nL][5]
 *   lam_out [nSeries][nTaus][nEmbs][nL][nThresholds][5]
 * Any output may be NULL to skip it (not all three).  Each (series, tau,
 * emb) block computes the requested measures from one pass over its pairs;
 * the series go up in one copy and the outputs come back in one copy each.
 * Results are those of the single-series calls, bit for bit.  Errors as
 * above; a bad argument for any series fails the whole call before any
 * work. */
void accrqa_amd_rqa_batch_f32(const float *series, const size_t *offsets, int nSeries, const int *tau_values,
                              int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nL,
                              const float *threshold_values, int nThresholds, int distance_type, float *rr_out,
                              float *det_out, float *lam_out, int *error);
void accrqa_amd_rqa_batch_f64(const double *series, const size_t *offsets, int nSeries, const int *tau_values,
                              int nTaus, const int *emb_values, int nEmbs, const int *lmin_values, int nL,
                              const double *threshold_values, int nThresholds, int distance_type, double *rr_out,
                              double *det_out, double *lam_out, int *error);

/* Options of the batched call (no counterpart in the reference).
 * autotune: 1 = the first call of each (precision, norm, threshold group)
 * whose first chunk has at least 2e9 pair decisions times the persistent
 * kernel's instantiations (different register budgets, so different waves
 * per SIMD) on that chunk and keeps the fastest for the process (default);
 * 0 = use the compiler's register budget.  graphs: 1 = the batched launches
 * of a call are captured in a hipGraph, replayed by later calls with the same
 * shapes, thresholds and device buffers (default); 0 = launched directly.
 * -1 leaves a setting unchanged.  Results are identical either way.  Returns
 * 0, or 2 for an out-of-range value. */
int accrqa_amd_batch_options(int autotune, int graphs);
/* Text report of the batched call's tuning choices (measured times per
 * instantiation) and graph builds and replays. */
const char *accrqa_amd_batch_report(void);

/* Recurrence plot.  Replaces the three overloads of accrqa_RP
 * (src/AccRQA_functions.cpp:996-1060, CPU only there:
 * rqa_CPU_R_matrix_parallel, src/AccRQA_CPU_function.cpp:226-244):
 * output[i*N_c + j] = 1 if the distance of rows i and j is within the
 * threshold ((threshold - distance) >= 0 in the input type), else 0, for
 * N_c = data_size - (emb - 1)*tau, row-major; char output for float and
 * double input, int output for double input, as in the reference.  The
 * output buffer holds N_c*N_c elements (accrqa_RP_output_size_in_elements
 * gives N_c).  Errors as above; the reference validates nothing here, this
 * returns 2 for tau < 0, emb < 1, N_c < 1 or a distance type other than 1/2. */
void accrqa_amd_RP_f32(char *output, const float *input_data, size_t data_size, int tau, int emb, float threshold,
                       int distance_type, int *error);
void accrqa_amd_RP_f64(char *output, const double *input_data, size_t data_size, int tau, int emb,
                       double threshold, int distance_type, int *error);
void accrqa_amd_RP_f64_i32(int *output, const double *input_data, size_t data_size, int tau, int emb,
                           double threshold, int distance_type, int *error);

/* The same with the series and the plot in device memory (N_c*N_c bytes at
 * d_output), launched on `stream` (the null stream when 0) and not
 * synchronised.  Returns 0 or an Accrqa_Error code. */
int accrqa_amd_rp_device_f32(const float *d_input, size_t data_size, int tau, int emb, float threshold,
                             int distance_type, char *d_output, void *stream);
int accrqa_amd_rp_device_f64(const double *d_input, size_t data_size, int tau, int emb, double threshold,
                             int distance_type, char *d_output, void *stream);

/* Form of the Euclidean distance (no counterpart in the reference, which has
 * one form per platform).  0 = automatic: the MFMA form for emb >= 8, exact
 * VALU sums otherwise; 1 = exact VALU sums only; 2 = MFMA form wherever it
 * applies (Euclidean norm; fp32 input, and fp64 input in the LAM pass, whose
 * Gram then runs on the series rounded to fp32).  The MFMA form computes
 * |x_i|^2 + |x_j|^2 - 2 x_i.x_j with the Gram tile on FP16 matrix cores
 * (each scaled sample split into two halves, hi/lo) and
 * re-decides every pair within a rigorous error bound of a threshold in the
 * reference's exact order and precision, so results are identical for every
 * form.  Process-wide; returns 0, or 2 for an unknown form. */
int accrqa_amd_set_form(int form);

/* LAM/DET work split (no counterpart in the reference).  0 = automatic: when
 * LAM is computed and N_c^2/2 bytes of rank tiles fit in under half of the
 * free device memory, the column pass ranks only the upper triangle and the
 * diagonal tiles and keeps them; the rest of each LAM column (row runs, R
 * being symmetric) and all of DET are then read from the tiles.  1 = never:
 * LAM ranks full columns and DET recomputes its pairs.  Results are identical
 * either way.  Process-wide; returns 0, or 2 for an unknown mode. */
int accrqa_amd_set_split(int mode);

/* Slab form of the split (no counterpart in the reference).  When the tile
 * store of all rows does not fit, the split form runs slab by slab, each slab
 * a band of rows whose tiles fit (LAM and DET together, one slab's tiles at a
 * time); slabs are as tall as fit.  rows > 0 (a multiple of 64, at least 256,
 * at most 2^29) takes the slab form whenever the split form applies, with slabs of that
 * many rows (for tests); 0 restores the automatic choice.  Results are
 * identical either way.  Returns 0 or ERR_INVALID_ARGUMENT. */
int accrqa_amd_set_slab_rows(int rows);

/* MFMA-form decision statistics (diagnostic, no counterpart in the
 * reference): accrqa_amd_mx_stats(1) zeroes the counters and switches them
 * on, 0 switches them off.  While on, every MFMA-form LAM pass that keeps the
 * rank tiles adds to counts[0] the pair elements it ranked (upper triangle and
 * the diagonal bands), counts[1] those whose table byte was flagged (a bound
 * within the Gram error bound of the bucket) and counts[2] those re-summed in
 * exact order; counts[2] / counts[0] is the recheck fraction.  A separate
 * kernel instantiation does the counting, so the default path is unchanged.
 * Both return 0 or an Accrqa_Error code. */
int accrqa_amd_mx_stats(int enable);
int accrqa_amd_mx_stats_read(unsigned long long *counts);

/* One matrix-core instruction of the MFMA form's Gram (diagnostic, no
 * counterpart in the reference), for checking the properties its error bound
 * assumes: d = c + a b with a = A[32][16] and b = B[16][32] as fp16 bit
 * patterns and c, d = [32][32] fp32, all row-major host arrays, computed by one
 * v_mfma_f32_32x32x16_f16 with the operands in the Gram kernels' lane layout.
 * Returns 0 or an Accrqa_Error code. */
int accrqa_amd_mfma16_probe(const uint16_t *a, const uint16_t *b, const float *c, float *d);

/* Recurrence plot, exact form (tuning aid, no counterpart in the reference):
 * 1 (the default) gives each lane four adjacent columns, so every store
 * writes 256 consecutive elements, where N_c % 4 == 0 and the plot's address
 * aligns those stores; 0 keeps one column per lane.  The plots are identical.
 * Returns 0 or ERR_INVALID_ARGUMENT. */
int accrqa_amd_set_rp_wide(int on);

/* Recurrence plot of bytes, both forms (tuning aid, no counterpart in the
 * reference): 1 (the default) decides the upper triangle only and stores
 * each 64 x 64 square mirrored as well, where N_c % 16 == 0 and the plot is
 * 16-byte aligned (every store is then 16 bytes of one row); 0 decides every
 * element.  The plots are identical.  Returns 0 or ERR_INVALID_ARGUMENT. */
int accrqa_amd_set_rp_sym(int on);

/* Waves per SIMD of the three tile scans (tuning aid, no counterpart in the
 * reference): at most `waves` workgroups of four waves each stay resident per
 * CU, held there by padding their LDS with dynamic LDS; 0 = as many as their
 * registers and LDS allow.  The default, 7, was measured faster than 8 on
 * this part.  Results are identical for every value.  Returns 0 or
 * ERR_INVALID_ARGUMENT (waves outside 0..8). */
int accrqa_amd_set_scan_waves(int waves);

/* Split form (no counterpart in the reference): 1 (the default) runs the DET
 * scan over the rank tiles on a second stream of the library's, beside the
 * LAM column and row scans on the call's stream, joined before the DET
 * measures are formed; 0 runs it after them.  Kernel profiling
 * (accrqa_amd_profile) runs them one after the other either way.  Results are
 * identical.  Returns 0 or ERR_INVALID_ARGUMENT. */
int accrqa_amd_set_overlap(int on);

/* Gram operand precision of the split MFMA form's LAM rank pass, for the
 * measured comparison of precisions (experiment, no counterpart in the
 * reference): 0 = FP16 hi/lo (the default and the only production form),
 * 1 = one FP16 half per sample, 2 = one BF16 half, 3 = one FP8 (E4M3) byte.
 * Every precision keeps every decision exact (its own error bound, then the
 * exact re-decision); the lower ones flag more pairs for it.  While it is not
 * 0, a pass that runs the split MFMA form needs fp32 input and at most 16
 * thresholds per block and returns ERR_INVALID_ARGUMENT otherwise.  Returns 0
 * or ERR_INVALID_ARGUMENT for an unknown value. */
int accrqa_amd_set_gram(int gram);

/* One matrix-core instruction of precision gram (1, 2 or 3 as above;
 * diagnostic): a = A[32][16], b = B[16][32] and c = C[32][32] fp32 row-major
 * host arrays; each operand is rounded as the Gram's window words are, qa and
 * qb receive the rounded values as the hardware decodes them, and
 * d = C + qa qb from one v_mfma_f32_32x32x16_{f16,bf16,fp8_fp8} in the Gram
 * kernels' lane layout.  Returns 0 or an Accrqa_Error code. */
int accrqa_amd_gram_probe(int gram, const float *a, const float *b, const float *c, float *d, float *qa,
                          float *qb);

/* Decisions in fp32 against true fp64 (diagnostic, no counterpart in the
 * reference): for fp32 series d_input (device memory) and each threshold,
 * mismatches[t] = the number of pairs (i, j) of the full N_c x N_c matrix
 * whose decision in the reference's fp32 arithmetic ((eps - dist) >= 0 with
 * dist summed in order in fp32) differs from the decision in fp64 arithmetic
 * on the same samples and threshold.  mismatches is a host array of
 * nThresholds entries; the call returns when they are written.  Returns 0 or
 * an Accrqa_Error code. */
int accrqa_amd_fp64_mismatch_f32(const float *d_input, size_t data_size, int tau, int emb,
                                 const float *threshold_values, int nThresholds, int distance_type,
                                 unsigned long long *mismatches, void *stream);

/* Kernel timing (tracing aid, no counterpart in the reference): when enabled,
 * every kernel launch is bracketed by HIP events on its own stream and the
 * elapsed times are summed per kernel name ("k_lam", "k_lam_mx",
 * "k_lam_vscan", "k_lam_hscan", "k_det", "k_det_mx", "k_det_scan", "k_rr",
 * "k_stitch", "k_finalize").  Adds a stream synchronisation per call. */
void accrqa_amd_profile(int enable);
void accrqa_amd_profile_reset(void);
/* Returns 0 and fills the totals if the kernel was seen, else 1 and zeros. */
int accrqa_amd_profile_read(const char *kernel, double *total_ms, long long *launches);

/* Name of the last device error seen by the library ("" if none). */
const char *accrqa_amd_last_error(void);

/* Library version "major.minor.patch". */
const char *accrqa_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif
