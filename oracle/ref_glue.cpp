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
#include "AccRQA_utilities_distance.hpp"
#include "AccRQA_utilities_error.hpp"

// The public RP overloads defined in src/AccRQA_functions.cpp:996-1060 (their
// header, include/AccRQA_functions.hpp, needs the build system's generated
// version macros, so the three signatures are declared here).
void accrqa_RP(char *output, float *input_data, size_t data_size, int tau, int emb, float threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);
void accrqa_RP(char *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);
void accrqa_RP(int *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);

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

// accrqa_RP, src/AccRQA_functions.cpp:996-1060 (the public API's three
// overloads: char output for float and double input, int output for double)
int ref_rp_u8_f32(char *out, float *x, long n, int tau, int emb, float thr, int dist) {
    Accrqa_Error e = SUCCESS;
    accrqa_RP(out, x, (size_t)n, tau, emb, thr, (Accrqa_Distance)dist, &e);
    return (int)e;
}
int ref_rp_u8_f64(char *out, double *x, long n, int tau, int emb, double thr, int dist) {
    Accrqa_Error e = SUCCESS;
    accrqa_RP(out, x, (size_t)n, tau, emb, thr, (Accrqa_Distance)dist, &e);
    return (int)e;
}
int ref_rp_i32_f64(int *out, double *x, long n, int tau, int emb, double thr, int dist) {
    Accrqa_Error e = SUCCESS;
    accrqa_RP(out, x, (size_t)n, tau, emb, thr, (Accrqa_Distance)dist, &e);
    return (int)e;
}

}  // extern "C"
