// AccRQA_amd.hpp -- the reference's public C++ API for RR / DET / LAM, served
// on the new AMD-GPU platform value.
//
// Names, overloads, argument order and meaning, output layouts and error
// codes follow include/AccRQA_functions.hpp:77-362 and the enums of
// include/AccRQA_utilities_{error,comp_platform,distance}.hpp, so code written
// against the reference compiles and links unchanged (the mangled symbols are
// identical).  The only addition is PLT_AMD_GPU = 2048 in Accrqa_CompPlatform.
//
// This library computes on the AMD GPU only.  Like the reference, which runs
// a platform it does not have on its default platform (silently for
// PLT_NV_GPU without CUDA, with a warning for unknown values;
// src/AccRQA_functions.cpp:525-539), it serves every platform value there:
// PLT_AMD_GPU and PLT_NV_GPU silently, PLT_CPU after a one-time note (the
// results equal the reference CPU path's bit for bit), unknown values after
// the reference's warning.
#ifndef ACCRQA_AMD_HPP
#define ACCRQA_AMD_HPP

#include <cstddef>
#include <cstdint>

enum Accrqa_Error {
    SUCCESS = 0,
    ERR_RUNTIME = 1,
    ERR_INVALID_ARGUMENT = 2,
    ERR_DATA_TYPE = 3,
    ERR_MEM_ALLOC_FAILURE = 4,
    ERR_MEM_COPY_FAILURE = 5,
    ERR_MEM_LOCATION = 6,
    ERR_CUDA_NOT_FOUND = 7,
    ERR_CUDA_DEVICE_NOT_FOUND = 8,
    ERR_CUDA_NOT_ENOUGH_MEMORY = 9,
    ERR_CUDA = 10,
    ERR_INVALID_METRIC_TYPE = 11
};
typedef enum Accrqa_Error Accrqa_Error;

enum Accrqa_CompPlatform {
    PLT_ERROR = 0,
    PLT_CPU = 1,
    PLT_NV_GPU = 1024,
    PLT_AMD_GPU = 2048
};
typedef enum Accrqa_CompPlatform Accrqa_CompPlatform;

enum Accrqa_Distance {
    DST_ERROR = 0,
    DST_EUCLIDEAN = 1,
    DST_MAXIMAL = 2
};
typedef enum Accrqa_Distance Accrqa_Distance;

void accrqa_print_error(Accrqa_Error *error);

void accrqa_LAM(float *output, float *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
                int nEmbs, int *vmin_values, int nVmins, float *threshold_values, int nThresholds,
                Accrqa_Distance distance_type, int calc_ENTR, Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
void accrqa_LAM(double *output, double *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
                int nEmbs, int *vmin_values, int nVmins, double *threshold_values, int nThresholds,
                Accrqa_Distance distance_type, int calc_ENTR, Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
int accrqa_LAM_output_size_in_elements(int nTaus, int nEmbs, int nVmins, int nThresholds);

void accrqa_DET(float *output, float *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
                int nEmbs, int *lmin_values, int nLmins, float *threshold_values, int nThresholds,
                Accrqa_Distance distance_type, int calc_ENTR, Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
void accrqa_DET(double *output, double *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
                int nEmbs, int *lmin_values, int nLmins, double *threshold_values, int nThresholds,
                Accrqa_Distance distance_type, int calc_ENTR, Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
int accrqa_DET_output_size_in_elements(int nTaus, int nEmbs, int nLmins, int nThresholds);

void accrqa_RR(float *output, float *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
               int nEmbs, float *threshold_values, int nThresholds, Accrqa_Distance distance_type,
               Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
void accrqa_RR(double *output, double *input_data, size_t data_size, int *tau_values, int nTaus, int *emb_values,
               int nEmbs, double *threshold_values, int nThresholds, Accrqa_Distance distance_type,
               Accrqa_CompPlatform comp_platform, Accrqa_Error *error);
int accrqa_RR_output_size_in_elements(int nTaus, int nEmbs, int nThresholds);

// Recurrence plot, the reference's three overloads (no platform argument:
// the reference computes it on the CPU, this library on the AMD GPU).
void accrqa_RP(char *output, float *input_data, size_t data_size, int tau, int emb, float threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);
void accrqa_RP(char *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);
void accrqa_RP(int *output, double *input_data, size_t data_size, int tau, int emb, double threshold,
               Accrqa_Distance distance_type, Accrqa_Error *error);
int64_t accrqa_RP_output_size_in_elements(int64_t input_size, int64_t tau, int64_t emb);

#endif
