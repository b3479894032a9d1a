// ref_glue.cpp -- exports raw line-length histograms of the REFERENCE CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference's own src/AccRQA_functions.cpp, straight from where the reference
// sources lie, into oracle/_ref/libaccrqa_ref.so (git-ignored).  The reference's
// public API returns only the five derived measures; these three wrappers call
// its internal histogram drivers so the golden fixtures can pin h_D and h_V
// bit for bit.  Nothing here re-implements reference code: each function is a
// one-line call into the reference template named in its comment.
#include "AccRQA_CPU_function.hpp"

extern "C" {

// rqa_CPU_DET_metric_parallel_mk1, src/AccRQA_CPU_function.cpp:587-642
void ref_det_hist_f32(unsigned long long *h, float *x, long n, float thr, int tau, int emb, int dist) {
    rqa_CPU_DET_metric_parallel_mk1<float>(h, thr, tau, emb, x, (int64_t)n, dist);
}
void ref_det_hist_f64(unsigned long long *h, double *x, long n, double thr, int tau, int emb, int dist) {
    rqa_CPU_DET_metric_parallel_mk1<double>(h, thr, tau, emb, x, (int64_t)n, dist);
}

// rqa_CPU_LAM_metric_parallel, src/AccRQA_CPU_function.cpp:446-490
void ref_lam_hist_f32(unsigned long long *h, float *x, long n, float thr, int tau, int emb, int dist) {
    rqa_CPU_LAM_metric_parallel<float>(h, thr, tau, emb, x, (int64_t)n, dist);
}
void ref_lam_hist_f64(unsigned long long *h, double *x, long n, double thr, int tau, int emb, int dist) {
    rqa_CPU_LAM_metric_parallel<double>(h, thr, tau, emb, x, (int64_t)n, dist);
}

// R_matrix_element, src/AccRQA_CPU_function.cpp:40-46
int ref_r_element_f32(float *x, long i, long j, float thr, int tau, int emb, int dist) {
    return R_matrix_element<float>(x, (size_t)i, (size_t)j, thr, tau, emb, dist);
}
int ref_r_element_f64(double *x, long i, long j, double thr, int tau, int emb, int dist) {
    return R_matrix_element<double>(x, (size_t)i, (size_t)j, thr, tau, emb, dist);
}

}  // extern "C"
