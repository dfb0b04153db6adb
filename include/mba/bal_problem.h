// bal_problem.h -- host C++ interface of the MI355X BA engine, mirroring slamplay's BA path.
//
//  * slamplay_amd::BALProblem keeps the accessor set and AoS layout of slamplay::BALProblem
//    (utils/datasets/BALProblem.h:31-119): 9 doubles per camera (angle-axis, t, f, k1, k2),
//    3 per point, observations x,y interleaved, cameras first then points in parameters().
//  * slamplay_amd::SolveBA(problem, options) replaces the file-local
//    `void SolveBA(BALProblem&)` of backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp:35,58.
//    It is a template so that it also accepts the reference's own slamplay::BALProblem unchanged:
//    any type with the same accessors maps onto ba_problem_view with zero copies.
#pragma once

#include <stdexcept>
#include <string>

#include "ba_c_api.h"

namespace slamplay_amd {

class BALProblem {
   public:
    /// BALProblem.cpp:78-139. Throws std::runtime_error when the file cannot be read (the
    /// reference prints and leaves the object uninitialised).
    explicit BALProblem(const std::string& filename, bool use_quaternions = false);
    /// Wraps a handle created by ba_synth_generate / ba_bal_from_arrays (takes ownership).
    explicit BALProblem(ba_bal* handle);
    ~BALProblem();
    BALProblem(const BALProblem&) = delete;
    BALProblem& operator=(const BALProblem&) = delete;

    void WriteToFile(const std::string& filename) const;      // BALProblem.cpp:141-182 (3-integer header)
    void WriteToPLYFile(const std::string& filename) const;   // BALProblem.cpp:185-228
    void Normalize();                                         // BALProblem.cpp:263-303
    void Perturb(double rotation_sigma, double translation_sigma, double point_sigma);  // :305-336

    int camera_block_size() const { return use_quaternions_ ? 10 : 9; }
    int point_block_size() const { return 3; }
    int num_cameras() const { return view_.num_cameras; }
    int num_points() const { return view_.num_points; }
    int num_observations() const { return view_.num_observations; }
    int num_parameters() const { return camera_block_size() * num_cameras() + 3 * num_points(); }
    const int* point_index() const { return view_.point_index; }
    const int* camera_index() const { return view_.camera_index; }
    const double* observations() const { return view_.observations; }
    const double* parameters() const { return view_.cameras; }
    const double* cameras() const { return view_.cameras; }
    const double* points() const { return view_.points; }
    double* mutable_cameras() { return view_.cameras; }
    double* mutable_points() { return view_.points; }
    double* mutable_camera_for_observation(int i) { return mutable_cameras() + camera_index()[i] * camera_block_size(); }
    double* mutable_point_for_observation(int i) { return mutable_points() + point_index()[i] * point_block_size(); }
    const double* camera_for_observation(int i) const { return cameras() + camera_index()[i] * camera_block_size(); }
    const double* point_for_observation(int i) const { return points() + point_index()[i] * point_block_size(); }

    ba_bal* handle() { return handle_; }

   private:
    ba_bal* handle_ = nullptr;
    ba_problem_view view_{};
    bool use_quaternions_ = false;
};

/// Solver options; defaults are Ceres Solver::Options defaults (what the reference leaves set).
struct SolveOptions : ba_options {
    SolveOptions() { ba_default_options(this); }
};

/// Solver summary with the per-iteration trace kept on the host; SolveBA records the problem
/// size and options so that FullReport() prints what ceres::Solver::Summary::FullReport() does
/// (test_bundle_adjustment_ceres.cpp:101).
struct SolveSummary : ba_summary {
    ba_iteration rows[512];
    int32_t num_cameras = 0, num_points = 0, num_observations = 0;
    ba_options options;
    SolveSummary() : ba_summary{} {
        trace = rows;
        trace_capacity = 512;
        ba_default_options(&options);
    }
    SolveSummary(const SolveSummary&) = delete;  // trace points into this object
    SolveSummary& operator=(const SolveSummary&) = delete;

    std::string FullReport() const {
        int64_t n = 0;
        if (ba_summary_full_report(this, num_cameras, num_points, num_observations, &options, nullptr, 0, &n) != BA_OK)
            throw std::runtime_error(ba_last_error_string());
        std::string r((size_t)n + 1, '\0');
        ba_summary_full_report(this, num_cameras, num_points, num_observations, &options, &r[0], n + 1, nullptr);
        r.resize((size_t)n);
        return r;
    }
};

/// Drop-in for SolveBA(BALProblem&): solves in place on the GPU. Returns ba_status.
template <typename Problem>
int SolveBA(Problem& bal_problem, const ba_options& options = SolveOptions(), ba_summary* summary = nullptr) {
    if (bal_problem.camera_block_size() != 9)
        throw std::invalid_argument("SolveBA: the Snavely functor needs 9-parameter cameras (no quaternions)");
    ba_problem_view v{};
    v.num_cameras = bal_problem.num_cameras();
    v.num_points = bal_problem.num_points();
    v.num_observations = bal_problem.num_observations();
    v.camera_block_size = 9;
    v.camera_index = bal_problem.camera_index();
    v.point_index = bal_problem.point_index();
    v.observations = bal_problem.observations();
    v.cameras = bal_problem.mutable_cameras();
    v.points = bal_problem.mutable_points();
    return ba_solve(&v, &options, summary);
}

/// SolveBA that also records what SolveSummary::FullReport() needs.
template <typename Problem>
int SolveBA(Problem& bal_problem, const ba_options& options, SolveSummary* summary) {
    if (summary) {
        summary->num_cameras = bal_problem.num_cameras();
        summary->num_points = bal_problem.num_points();
        summary->num_observations = bal_problem.num_observations();
        summary->options = options;
    }
    return SolveBA(bal_problem, options, static_cast<ba_summary*>(summary));
}

}  // namespace slamplay_amd
