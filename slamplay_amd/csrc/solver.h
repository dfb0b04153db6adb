// solver.h -- host-side declarations shared by the single-problem solver sources.
#pragma once

#include <memory>
#include <vector>

#include "common.h"
#include "stereo.h"

namespace mba {

// Device layout built on the host once per problem (see lm_kernels.h for its meaning).
struct Layout {
    int nc = 0, np = 0, no = 0, n_chunks = 0;
    std::vector<int> pt_perm;    // internal point -> caller point
    std::vector<int> co_orig;    // camera-order observation -> caller observation
    std::vector<int> cam_off;    // nc + 1
    std::vector<int> co_pt;      // internal point of each camera-order observation
    std::vector<double> co_uv;   // x, y per camera-order observation
    std::vector<int> chunk_off, chunk_len;
    std::vector<int> slot_of;    // camera-order observation -> SELL slot
    std::vector<int> slot_j;     // SELL slot -> camera-order observation (-1: padding)
};

int build_layout(const ba_problem_view* v, Layout* L);

// 9x9 blocks (a <= b) of the reduced camera system and, per block, the ordered observation pairs
// (i, j) of a common point with camera(i) = a, camera(j) = b (camera-order indices).  Every
// diagonal block is present, possibly with no pairs.
struct PairBlocks {
    std::vector<int> a, b, off;  // off: nblocks + 1
    std::vector<int> i, j;
};
void build_pair_blocks(const Layout& L, PairBlocks* out);

struct SolverBase {
    virtual ~SolverBase() = default;
    virtual int run(ba_summary* summary) = 0;
    virtual int reset() = 0;
    virtual int download(double* cameras, double* points) = 0;
    // BALProblem::Normalize / ::Perturb on the device-held parameters (preprocess.h)
    virtual int normalize() = 0;
    virtual int perturb(uint32_t seed, double rotation_sigma, double translation_sigma, double point_sigma) = 0;
    virtual int set_profiling(bool on) = 0;
    virtual int kernel_times(ba_kernel_time* out, int cap, int* count) = 0;
    virtual int64_t device_bytes() const = 0;
    virtual int solver_kind() const = 0;
    // > 0: run()'s waits for the device give up with BA_ERR_TIMEOUT after this many seconds.
    virtual void set_wait_timeout(double seconds) = 0;
};

// The full_slam stereo model (stereo.h): rig and per-observation camera side (caller order).
struct StereoInput {
    const StereoRig* rig;
    const int32_t* side;
};

int make_solver(const ba_problem_view* v, const ba_options& o, std::unique_ptr<SolverBase>* out,
                const StereoInput* stereo = nullptr);

// A full_slam stereo window in the solver's terms: poses in 9-parameter camera slots, the rig, and
// the problem view over them (its points are the caller's).
struct StereoProblem {
    std::vector<double> cams;
    StereoRig rig;
    ba_problem_view view;
    StereoInput input;
};
// Validates a stereo view and fills sp (the view's pose / rig conventions of ba_stereo_solve).
int stereo_prepare(const ba_stereo_view* v, StereoProblem* sp);
// Writes the solved poses (sp.cams) back into v and applies the reference's outlier rule
// (full_slam/src/backend.cpp:153-186) from chi2_start to the residuals of v's points.
void stereo_finish(const ba_stereo_view* v, const StereoProblem& sp, double chi2_start, uint8_t* outlier,
                   double* chi2_threshold);

// BA_OK, or BA_ERR_INVALID_ARGUMENT with the reason for an option out of range.
int validate_options(const ba_options* o);
// The device option block of ba_options (g2o mode: lambda-LM, no column scaling).
struct Opts;
Opts solver_opts(const ba_options& o);

}  // namespace mba
