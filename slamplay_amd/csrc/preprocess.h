// preprocess.h -- Normalize and Perturb of a device-resident problem (BALProblem.cpp:263-336).
#pragma once

#include <cstdint>

#include "common.h"

namespace mba {

// Parameters as the single-problem solver holds them: cameras SoA (cam[i * nc + c], 9 rows),
// points 4 doubles each in the solver's internal order, pt_orig[ip] = caller index of point ip.
struct DeviceParams {
    double* cam;
    double* pt;
    const int* pt_orig;
    int nc, np;
};

// Scratch owned by the caller's handle, grown on demand and kept between calls.
struct PrepScratch {
    DevBuf<unsigned long long> sel_state;  // 4 selections x (prefix, mask, rank)
    DevBuf<unsigned int> hist;             // digit counts of the four selections (256 each)
    DevBuf<double> dist;                   // L1 distances to the median
    DevBuf<unsigned int> draws;            // glibc rand() outputs
    DevBuf<unsigned int> tables;           // jump-ahead matrices
    DevBuf<unsigned int> counts;           // accepted pairs per block, then their exclusive scan
    DevBuf<double> normals;
    DevBuf<int> total;                     // accepted pairs of the whole stream
    int table_chunks = 0;
};

// BALProblem::Normalize (BALProblem.cpp:263-303) on the device: per-axis upper medians and the
// upper median of the L1 distances by radix selection, then the point and camera-centre transform.
int device_normalize(const DeviceParams& p, PrepScratch* w, hipStream_t s);

// BALProblem::Perturb (BALProblem.cpp:305-336) on the device with the normals glibc's rand() gives
// after srand(seed): the rand() stream is generated in parallel by jump-ahead of its additive
// lagged-Fibonacci recurrence, the polar method's accepted pairs are compacted by a prefix sum, and
// the k-th normal goes where the reference's sequential loop would put it.
int device_perturb(const DeviceParams& p, uint32_t seed, double rotation_sigma, double translation_sigma,
                   double point_sigma, PrepScratch* w, hipStream_t s);

}  // namespace mba
