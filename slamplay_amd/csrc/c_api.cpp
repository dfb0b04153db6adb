// c_api.cpp -- error reporting, defaults and argument validation of the C ABI.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "common.h"

namespace mba {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

namespace {
std::mutex g_lane_mu;
std::map<int, std::vector<hipStream_t>> g_lane_free;  // per device: pooled lane streams not in use
std::map<int, bool> g_lane_made;
}  // namespace

int use_device(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(BA_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(BA_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, n);
    BA_HIP(hipSetDevice(device));
    std::lock_guard<std::mutex> lk(g_lane_mu);
    if (!g_lane_made[device]) {
        std::vector<hipStream_t>& pool = g_lane_free[device];
        for (int i = 0; i < kLaneStreams; ++i) {
            hipStream_t st = nullptr;
            BA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            pool.push_back(st);
        }
        g_lane_made[device] = true;
    }
    return BA_OK;
}

int acquire_lane_streams(int device, int count, std::vector<hipStream_t>* out, std::vector<bool>* pooled) {
    out->clear();
    pooled->clear();
    {
        std::lock_guard<std::mutex> lk(g_lane_mu);
        std::vector<hipStream_t>& pool = g_lane_free[device];
        while ((int)out->size() < count && !pool.empty()) {
            out->push_back(pool.front());
            pool.erase(pool.begin());
            pooled->push_back(true);
        }
    }
    while ((int)out->size() < count) {
        hipStream_t st = nullptr;
        BA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        out->push_back(st);
        pooled->push_back(false);
    }
    return BA_OK;
}

void release_lane_streams(int device, const std::vector<hipStream_t>& streams, const std::vector<bool>& pooled) {
    std::lock_guard<std::mutex> lk(g_lane_mu);
    std::vector<hipStream_t>& pool = g_lane_free[device];
    for (size_t i = 0; i < streams.size(); ++i) {
        if (pooled[i])
            pool.push_back(streams[i]);
        else
            (void)hipStreamDestroy(streams[i]);
    }
}

int wait_stream(hipStream_t s, hipEvent_t ev, double timeout_s, const char* what) {
    if (!(timeout_s > 0.0) || !ev) {
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }
    BA_HIP(hipEventRecord(ev, s));
    const auto t0 = std::chrono::steady_clock::now();
    for (int polls = 0;; ++polls) {
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) return BA_OK;
        if (e != hipErrorNotReady)
            return fail(e == hipErrorOutOfMemory ? BA_ERR_OUT_OF_MEMORY : BA_ERR_HIP, "%s: %s", what,
                        hipGetErrorString(e));
        const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (waited > timeout_s)
            return fail(BA_ERR_TIMEOUT, "%s: device work not finished after %.1f s", what, waited);
        // spin briefly (most waits are short), then yield the core
        if (polls > 64) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int allow_dynamic_lds(const void* kernel, int bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, int> granted;  // (device, kernel) -> bytes
    int dev = 0;
    BA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    int& have = granted[{dev, kernel}];
    if (have >= bytes) return BA_OK;
    BA_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    have = bytes;
    return BA_OK;
}

int validate_view(const ba_problem_view* v, bool need_params) {
    if (!v) return fail(BA_ERR_INVALID_ARGUMENT, "problem view is NULL");
    if (v->camera_block_size != 9)
        return fail(BA_ERR_INVALID_ARGUMENT, "camera_block_size must be 9 (got %d; quaternion cameras are not "
                    "supported by the Snavely residual)", v->camera_block_size);
    if (v->num_cameras < 0 || v->num_points < 0 || v->num_observations < 0)
        return fail(BA_ERR_INVALID_ARGUMENT, "negative problem size");
    if (v->num_observations > 0 && (!v->camera_index || !v->point_index || !v->observations))
        return fail(BA_ERR_INVALID_ARGUMENT, "observation arrays are NULL");
    if (need_params && ((v->num_cameras > 0 && !v->cameras) || (v->num_points > 0 && !v->points)))
        return fail(BA_ERR_INVALID_ARGUMENT, "parameter arrays are NULL");
    for (int k = 0; k < v->num_observations; ++k) {
        const int c = v->camera_index[k], p = v->point_index[k];
        if (c < 0 || c >= v->num_cameras)
            return fail(BA_ERR_INVALID_ARGUMENT, "observation %d: camera index %d out of range [0,%d)", k, c,
                        v->num_cameras);
        if (p < 0 || p >= v->num_points)
            return fail(BA_ERR_INVALID_ARGUMENT, "observation %d: point index %d out of range [0,%d)", k, p,
                        v->num_points);
        if (!std::isfinite(v->observations[2 * (size_t)k]) || !std::isfinite(v->observations[2 * (size_t)k + 1]))
            return fail(BA_ERR_NON_FINITE, "observation %d is not finite", k);
    }
    if (need_params) {
        for (size_t i = 0; i < 9 * (size_t)v->num_cameras; ++i)
            if (!std::isfinite(v->cameras[i]))
                return fail(BA_ERR_NON_FINITE, "camera parameter %zu is not finite", i);
        for (size_t i = 0; i < 3 * (size_t)v->num_points; ++i)
            if (!std::isfinite(v->points[i])) return fail(BA_ERR_NON_FINITE, "point parameter %zu is not finite", i);
    }
    return BA_OK;
}

}  // namespace mba

extern "C" {

int32_t ba_abi_version(void) { return BA_ABI_VERSION; }

const char* ba_last_error_string(void) { return mba::g_last_error.c_str(); }

void ba_default_options(ba_options* o) {
    if (!o) return;
    *o = ba_options{};
    o->max_num_iterations = 50;
    o->max_num_consecutive_invalid_steps = 5;
    o->jacobi_scaling = 1;
    o->loss = BA_LOSS_HUBER;
    o->linear_solver = BA_SOLVER_AUTO;
    o->pcg_max_iterations = 50;
    o->precision = BA_PRECISION_FP64;
    o->dense_schur_max_cameras = 300;
    o->device = 0;
    o->verbose = 0;
    o->refine_steps = 2;
    o->loss_scale = 1.0;
    o->function_tolerance = 1e-6;
    o->gradient_tolerance = 1e-10;
    o->parameter_tolerance = 1e-8;
    o->initial_trust_region_radius = 1e4;
    o->max_trust_region_radius = 1e16;
    o->min_trust_region_radius = 1e-32;
    o->min_relative_decrease = 1e-3;
    o->min_lm_diagonal = 1e-6;
    o->max_lm_diagonal = 1e32;
    o->pcg_tolerance = 1e-6;
    o->mode = BA_MODE_CERES;
    o->g2o_max_trials = 10;
    o->g2o_tau = 1e-5;
}

int ba_device_count(int32_t* count) {
    mba::clear_error();
    if (!count) return mba::fail(BA_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return BA_OK;
}

}  // extern "C"
