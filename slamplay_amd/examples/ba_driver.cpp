// ba_driver.cpp -- the Ceres driver of backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp
// (main at :37-56, SolveBA at :58-101) with SolveBA replaced by the engine's drop-in.
//
//   ba_driver <bal_file> [out_dir] [g2o]
//
// Loads the BAL file, Normalize(), Perturb(0.1, 0.5, 0.5) on glibc rand() (this process never calls
// srand, as the reference's driver), writes initial.ply, solves on the GPU in place, writes
// final.ply, prints the summary's FullReport() as the reference does after ceres::Solve (:101), then
// one line: "initial_cost final_cost iterations termination".  With "g2o" the
// solve uses the g2o driver's rules and its optimize(40) (test_bundle_adjustment_g2o.cpp:166-218).
#include <cstdio>
#include <iostream>
#include <string>

#include "mba/bal_problem.h"

namespace {

// The reference's entry point, kept file-local as there (test_bundle_adjustment_ceres.cpp:35,58).
int SolveBA(slamplay_amd::BALProblem& bal_problem, bool g2o, slamplay_amd::SolveSummary* summary) {
    slamplay_amd::SolveOptions options;  // Ceres Solver::Options defaults, BA_SOLVER_AUTO
    if (g2o) {
        options.mode = BA_MODE_G2O;
        options.max_num_iterations = 40;
    }
    return slamplay_amd::SolveBA(bal_problem, options, summary);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || argc > 4) {
        std::cerr << "usage: " << argv[0] << " bal_data.txt [out_dir] [g2o]" << std::endl;
        return 2;
    }
    const std::string out_dir = argc >= 3 ? std::string(argv[2]) + "/" : std::string();
    const bool g2o = argc == 4 && std::string(argv[3]) == "g2o";
    try {
        slamplay_amd::BALProblem bal_problem(argv[1]);
        bal_problem.Normalize();
        bal_problem.Perturb(0.1, 0.5, 0.5);
        bal_problem.WriteToPLYFile(out_dir + "initial.ply");
        slamplay_amd::SolveSummary summary;
        const int st = SolveBA(bal_problem, g2o, &summary);
        if (st != BA_OK) {
            std::cerr << "SolveBA failed: " << ba_last_error_string() << std::endl;
            return 1;
        }
        bal_problem.WriteToPLYFile(out_dir + "final.ply");
        std::cout << summary.FullReport() << "\n" << std::flush;
        std::printf("%.17g %.17g %d %d\n", summary.initial_cost, summary.final_cost, summary.num_iterations,
                    summary.termination_type);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
