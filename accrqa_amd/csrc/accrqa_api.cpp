// accrqa_api.cpp -- the reference's C++ entry points (include/AccRQA_amd.hpp)
// on top of the C ABI (include/accrqa_amd.h).
//
// Dispatch mirrors accrqa_RR / accrqa_DET / accrqa_LAM
// (src/AccRQA_functions.cpp:525-555, 780-810, 945-975) with the GPU branch
// pointed at the gfx950 engine.  Argument checks and their order (counts
// first -> ERR_INVALID_ARGUMENT, then NULL arrays -> ERR_MEM_ALLOC_FAILURE)
// are the reference's (AccRQA_functions.cpp:486-491, 741-746, 895-900).
#include <cstdio>

#include "AccRQA_amd.hpp"
#include "accrqa_amd.h"

namespace {

// Platform dispatch of this library.  The reference runs PLT_NV_GPU without
// CUDA on its default platform silently, and any value other than PLT_CPU /
// PLT_NV_GPU on its default platform after a warning
// (src/AccRQA_functions.cpp:525-539).  Here the only platform is the AMD GPU,
// whose results equal PLT_CPU's bit for bit, so every value is served there:
// PLT_AMD_GPU and PLT_NV_GPU silently, PLT_CPU with a one-time note (the
// caller asked for the host), other values with the reference's warning.
void serve_platform(Accrqa_CompPlatform p) {
    static bool cpu_noted = false;
    if (p == PLT_AMD_GPU || p == PLT_NV_GPU) return;
    if (p == PLT_CPU) {
        if (!cpu_noted)
            std::printf("NOTE: libaccrqa_amd computes PLT_CPU requests on the AMD GPU (same results as the CPU path).\n");
        cpu_noted = true;
        return;
    }
    std::printf("WARNING: Unknown compute platform. Defaulting to AMD GPU.\n");
}

}  // namespace

void accrqa_print_error(Accrqa_Error *error) {
    const char *msg = "unrecognised AccRQA error";
    switch (*error) {
        case SUCCESS: msg = "success"; break;
        case ERR_RUNTIME: msg = "generic runtime error"; break;
        case ERR_INVALID_ARGUMENT: msg = "wrong arguments"; break;
        case ERR_DATA_TYPE: msg = "unsupported data type"; break;
        case ERR_MEM_ALLOC_FAILURE: msg = "array not allocated"; break;
        case ERR_MEM_COPY_FAILURE: msg = "memory copy failure host<->device"; break;
        case ERR_MEM_LOCATION: msg = "wrong memory location"; break;
        case ERR_CUDA_NOT_FOUND: msg = "GPU runtime not found"; break;
        case ERR_CUDA_DEVICE_NOT_FOUND: msg = "could not locate GPU device"; break;
        case ERR_CUDA_NOT_ENOUGH_MEMORY: msg = "not enough memory on the device (GPU)"; break;
        case ERR_CUDA: msg = "other GPU error"; break;
        case ERR_INVALID_METRIC_TYPE: msg = "invalid metric selected"; break;
    }
    std::printf("%s", msg);
}

#define ACCRQA_DL(NAME, T, SFX)                                                                                   \
    void NAME(T *output, T *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values, int nEmbs, \
              int *l_values, int nL, T *threshold_values, int nThresholds, Accrqa_Distance distance_type,        \
              int calc_ENTR, Accrqa_CompPlatform comp_platform, Accrqa_Error *error) {                           \
        serve_platform(comp_platform);                                                                           \
        int e = 0;                                                                                               \
        NAME##_amd_##SFX(output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, l_values, nL,      \
                         threshold_values, nThresholds, (int)distance_type, calc_ENTR, &e);                      \
        if (error) *error = (Accrqa_Error)e;                                                                     \
    }

// accrqa_amd_DET_f32 etc. are reached through these aliases so the macro can
// paste the C ABI name from the C++ one.
#define accrqa_DET_amd_f32 accrqa_amd_DET_f32
#define accrqa_DET_amd_f64 accrqa_amd_DET_f64
#define accrqa_LAM_amd_f32 accrqa_amd_LAM_f32
#define accrqa_LAM_amd_f64 accrqa_amd_LAM_f64

ACCRQA_DL(accrqa_DET, float, f32)
ACCRQA_DL(accrqa_DET, double, f64)
ACCRQA_DL(accrqa_LAM, float, f32)
ACCRQA_DL(accrqa_LAM, double, f64)

void accrqa_RR(float *output, float *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
               int nEmbs, float *threshold_values, int nThresholds, Accrqa_Distance distance_type,
               Accrqa_CompPlatform comp_platform, Accrqa_Error *error) {
    serve_platform(comp_platform);
    int e = 0;
    accrqa_amd_RR_f32(output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, threshold_values,
                      nThresholds, (int)distance_type, &e);
    if (error) *error = (Accrqa_Error)e;
}

void accrqa_RR(double *output, double *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
               int nEmbs, double *threshold_values, int nThresholds, Accrqa_Distance distance_type,
               Accrqa_CompPlatform comp_platform, Accrqa_Error *error) {
    serve_platform(comp_platform);
    int e = 0;
    accrqa_amd_RR_f64(output, input_data, data_size, tau_values, nTaus, emb_values, nEmbs, threshold_values,
                      nThresholds, (int)distance_type, &e);
    if (error) *error = (Accrqa_Error)e;
}

// accrqa_RP (AccRQA_functions.cpp:996-1060): no platform argument there (the
// reference computes it on the CPU); here it runs on the AMD GPU.  As in the
// reference, a call whose *error is not SUCCESS on entry does nothing.
void accrqa_RP(char *output, float *input_data, size_t data_size, int tau, int emb, float threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error) {
    if (*error != SUCCESS) return;
    int e = 0;
    accrqa_amd_RP_f32(output, input_data, data_size, tau, emb, threshold, (int)distance_type, &e);
    *error = (Accrqa_Error)e;
}
void accrqa_RP(char *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error) {
    if (*error != SUCCESS) return;
    int e = 0;
    accrqa_amd_RP_f64(output, input_data, data_size, tau, emb, threshold, (int)distance_type, &e);
    *error = (Accrqa_Error)e;
}
void accrqa_RP(int *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error) {
    if (*error != SUCCESS) return;
    int e = 0;
    accrqa_amd_RP_f64_i32(output, input_data, data_size, tau, emb, threshold, (int)distance_type, &e);
    *error = (Accrqa_Error)e;
}
// AccRQA_functions.cpp:989-994: N_c (the plot is N_c x N_c)
int64_t accrqa_RP_output_size_in_elements(int64_t input_size, int64_t tau, int64_t emb) {
    return input_size - (emb - 1) * tau;
}

// Output sizes (AccRQA_functions.cpp:558-560, 812-814, 978-982).
int accrqa_LAM_output_size_in_elements(int nTaus, int nEmbs, int nVmins, int nThresholds) {
    return nTaus * nEmbs * nVmins * nThresholds * 5;
}
int accrqa_DET_output_size_in_elements(int nTaus, int nEmbs, int nLmins, int nThresholds) {
    return nTaus * nEmbs * nLmins * nThresholds * 5;
}
int accrqa_RR_output_size_in_elements(int nTaus, int nEmbs, int nThresholds) { return nTaus * nEmbs * nThresholds; }
