// report.cpp -- the text of Ceres's Solver::Summary::FullReport() for a solve of the engine.
//
// The reference's Ceres driver prints summary.FullReport() after ceres::Solve
// (backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp:101).  The layout (problem
// size table, minimizer, linear solver given/used, costs, step counts, times, termination line) is
// kept because users read and grep it; the figures come from ba_summary.  Figures Ceres has and the
// engine does not (per-phase evaluation times, threads, the elimination ordering's group sizes) are
// reported as the engine's equivalents: one device, the point/camera elimination groups of the
// Schur complement and the device time of the minimizer from the trace.
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "common.h"

namespace mba {
namespace {

void add(std::string* s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    *s += buf;
}

const char* solver_name(int s) {
    switch (s) {
        case BA_SOLVER_DENSE_SCHUR: return "DENSE_SCHUR";
        case BA_SOLVER_ITERATIVE_SCHUR: return "ITERATIVE_SCHUR";
        default: return "SPARSE_SCHUR";  // BA_SOLVER_AUTO stands in for the driver's SPARSE_SCHUR
    }
}

const char* termination_name(int t) {
    switch (t) {
        case BA_CONVERGENCE: return "CONVERGENCE";
        case BA_NO_CONVERGENCE: return "NO_CONVERGENCE";
        case BA_USER_FAILURE: return "USER_FAILURE";
        default: return "FAILURE";
    }
}

// The termination message: which test fired and the value it compared against its option.
std::string termination_message(const ba_summary& sm, const ba_options& o) {
    std::string m;
    const double v = sm.termination_value;
    switch (sm.termination_reason) {
        case BA_REASON_MAX_ITERATIONS:
            add(&m, "Maximum number of iterations reached. Number of iterations: %d.", (int)v);
            break;
        case BA_REASON_FUNCTION_TOLERANCE:
            add(&m, "Function tolerance reached. |cost_change|/cost: %e <= %e", v, o.function_tolerance);
            break;
        case BA_REASON_GRADIENT_TOLERANCE:
            add(&m, "Gradient tolerance reached. Gradient max norm: %e <= %e", v, o.gradient_tolerance);
            break;
        case BA_REASON_PARAMETER_TOLERANCE:
            add(&m, "Parameter tolerance reached. Relative step_norm: %e <= %e.", v, o.parameter_tolerance);
            break;
        case BA_REASON_MIN_RADIUS:
            add(&m, "Minimum trust region radius reached. Trust region radius: %e <= %e", v,
                o.min_trust_region_radius);
            break;
        case BA_REASON_INVALID_STEPS:
            add(&m, "Number of consecutive invalid steps more than Solver::Options::max_num_consecutive_invalid_steps: %d",
                o.max_num_consecutive_invalid_steps);
            break;
        case BA_REASON_NON_FINITE_COST:
            add(&m, "Residual and Jacobian evaluation failed: the cost is not finite.");
            break;
        case BA_REASON_G2O_STOP:
            add(&m, "g2o Levenberg stopped: trial budget of %d exhausted or no decrease (rho = %e).",
                o.g2o_max_trials, v);
            break;
        default:
            if (sm.termination_type == BA_CONVERGENCE)
                add(&m, "Function tolerance reached. No non-constant parameter blocks found.");
            else
                add(&m, "Solve ended before a termination test fired.");
    }
    return m;
}

}  // namespace

std::string full_report(const ba_summary& sm, int nc, int np, int no, const ba_options& o) {
    std::string s;
    add(&s, "\nSolver Summary (slamplay_amd %d, HIP gfx950)\n\n", BA_ABI_VERSION);
    add(&s, "%-25s%16s%25s\n", "", "Original", "Reduced");
    const int64_t blocks = (int64_t)nc + np, params = 9 * (int64_t)nc + 3 * (int64_t)np;
    add(&s, "%-25s%16" PRId64 "%25" PRId64 "\n", "Parameter blocks", blocks, blocks);
    add(&s, "%-25s%16" PRId64 "%25" PRId64 "\n", "Parameters", params, params);
    add(&s, "%-25s%16d%25d\n", "Residual blocks", no, no);
    add(&s, "%-25s%16" PRId64 "%25" PRId64 "\n\n", "Residuals", 2 * (int64_t)no, 2 * (int64_t)no);
    add(&s, "%-25s%16s\n\n", "Minimizer", "TRUST_REGION");
    add(&s, "%-25s%16s\n", "Dense linear algebra", "HIP_GFX950");
    add(&s, "%-25s%16s\n\n", "Trust region strategy", o.mode == BA_MODE_G2O ? "G2O_LEVENBERG" : "LEVENBERG_MARQUARDT");
    add(&s, "%-25s%16s%25s\n", "", "Given", "Used");
    add(&s, "%-25s%16s%25s\n", "Linear solver", solver_name(o.linear_solver), solver_name(sm.linear_solver_used));
    if (sm.linear_solver_used == BA_SOLVER_ITERATIVE_SCHUR)
        add(&s, "%-25s%16s%25s\n", "Preconditioner", "SCHUR_JACOBI", "SCHUR_JACOBI");
    add(&s, "%-25s%16s%25s\n", "Precision", o.precision == BA_PRECISION_MIXED ? "MIXED" : "FP64",
        o.precision == BA_PRECISION_MIXED ? "MIXED" : "FP64");
    add(&s, "%-25s%16d%25d\n", "Devices", 1, 1);
    char order[64];
    snprintf(order, sizeof(order), "%d,%d", np, nc);
    add(&s, "%-25s%16s%25s\n", "Linear solver ordering", "AUTOMATIC", order);
    add(&s, "%-25s%16s%25s\n\n", "Schur structure", "2,3,9", "2,3,9");
    add(&s, "Cost:\n");
    add(&s, "%-25s%16e\n", "Initial", sm.initial_cost);
    if (sm.termination_type != BA_FAILURE) {
        add(&s, "%-25s%16e\n", "Final", sm.final_cost);
        add(&s, "%-25s%16e\n", "Change", sm.initial_cost - sm.final_cost);
    }
    add(&s, "\n%-25s%16d\n", "Minimizer iterations", sm.num_successful_steps + sm.num_unsuccessful_steps);
    add(&s, "%-25s%16d\n", "Successful steps", sm.num_successful_steps);
    add(&s, "%-25s%16d\n", "Unsuccessful steps", sm.num_unsuccessful_steps);
    if (sm.linear_solver_used == BA_SOLVER_ITERATIVE_SCHUR)
        add(&s, "%-25s%16d\n", "Linear solver iterations", sm.total_linear_solver_iterations);
    add(&s, "\nTime (in seconds):\n");
    double minimizer = 0.0;
    if (sm.trace && sm.trace_capacity > 0 && sm.num_iterations > 0) {
        const int last = (sm.num_iterations < sm.trace_capacity ? sm.num_iterations : sm.trace_capacity) - 1;
        minimizer = sm.trace[last].cumulative_time_in_seconds;
    }
    const double total = sm.total_time_ms * 1e-3;
    if (minimizer > 0.0) {
        add(&s, "%-25s%16.6f\n", "Preprocessor", total > minimizer ? total - minimizer : 0.0);
        add(&s, "%-25s%16.6f\n", "Minimizer (device)", minimizer);
    }
    add(&s, "%-25s%16.6f\n\n", "Total", total);
    add(&s, "%-25s%16s (%s)\n", "Termination:", termination_name(sm.termination_type),
        termination_message(sm, o).c_str());
    return s;
}

}  // namespace mba

extern "C" int ba_summary_full_report(const ba_summary* summary, int32_t num_cameras, int32_t num_points,
                                      int32_t num_observations, const ba_options* options, char* out,
                                      int64_t capacity, int64_t* length) {
    mba::clear_error();
    if (!summary) return mba::fail(BA_ERR_INVALID_ARGUMENT, "summary is NULL");
    if (num_cameras < 0 || num_points < 0 || num_observations < 0)
        return mba::fail(BA_ERR_INVALID_ARGUMENT, "negative problem size");
    if (capacity < 0 || (capacity > 0 && !out)) return mba::fail(BA_ERR_INVALID_ARGUMENT, "output buffer is NULL");
    ba_options o;
    if (options)
        o = *options;
    else
        ba_default_options(&o);
    const std::string r = mba::full_report(*summary, num_cameras, num_points, num_observations, o);
    if (length) *length = (int64_t)r.size();
    if (capacity > 0) {
        const size_t n = r.size() < (size_t)capacity ? r.size() : (size_t)capacity - 1;
        memcpy(out, r.data(), n);
        out[n] = '\0';
    }
    return BA_OK;
}
