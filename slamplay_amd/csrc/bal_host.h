// bal_host.h -- host-side BAL problem container behind the ba_bal handle.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/mba/ba_c_api.h"

struct ba_bal {
    int nc = 0, np = 0, no = 0;
    bool use_quaternions = false;
    std::vector<int> ci, pi;       // camera / point index per observation
    std::vector<double> obs;       // x, y per observation
    std::vector<double> params;    // cameras (9 or 10 each) then points (3 each)

    int camera_block_size() const { return use_quaternions ? 10 : 9; }
    double* points() { return params.data() + (size_t)camera_block_size() * nc; }
    const double* points() const { return params.data() + (size_t)camera_block_size() * nc; }
    ba_problem_view view() {
        ba_problem_view v{};
        v.num_cameras = nc;
        v.num_points = np;
        v.num_observations = no;
        v.camera_block_size = camera_block_size();
        v.camera_index = ci.data();
        v.point_index = pi.data();
        v.observations = obs.data();
        v.cameras = params.data();
        v.points = points();
        return v;
    }
};

namespace mba {
void host_rotate_point(const double aa[3], const double pt[3], double out[3]);
void host_camera_to_aa_center(const double* cam, bool quat, double aa[3], double center[3]);
void host_aa_center_to_camera(const double aa[3], const double center[3], bool quat, double* cam);
int bal_load(const char* path, bool use_quaternions, ba_bal** out);
int bal_write(const ba_bal* b, const char* path, bool legacy_header);
int bal_write_ply(const ba_bal* b, const char* path);
void bal_normalize(ba_bal* b);
void bal_perturb(ba_bal* b, double rotation_sigma, double translation_sigma, double point_sigma);
}  // namespace mba
