// bal_problem.cpp -- host BALProblem: BAL text I/O, Normalize, Perturb, PLY export
// (utils/datasets/BALProblem.{h,cpp} of slamplay), exposed through ba_bal handles and the
// slamplay_amd::BALProblem class.
#include "bal_host.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>

#include "../../include/mba/bal_problem.h"
#include "common.h"

namespace mba {

// rotation.h:133-198 in double (Rodrigues with the DBL_EPSILON first-order branch).
void host_rotate_point(const double aa[3], const double pt[3], double out[3]) {
    const double theta2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
    if (theta2 > std::numeric_limits<double>::epsilon()) {
        const double theta = std::sqrt(theta2);
        const double c = std::cos(theta), s = std::sin(theta);
        const double inv = 1.0 / theta;
        const double w[3] = {aa[0] * inv, aa[1] * inv, aa[2] * inv};
        const double wx[3] = {w[1] * pt[2] - w[2] * pt[1], w[2] * pt[0] - w[0] * pt[2], w[0] * pt[1] - w[1] * pt[0]};
        const double tmp = (w[0] * pt[0] + w[1] * pt[1] + w[2] * pt[2]) * (1.0 - c);
        for (int i = 0; i < 3; ++i) out[i] = pt[i] * c + wx[i] * s + w[i] * tmp;
    } else {
        const double wx[3] = {aa[1] * pt[2] - aa[2] * pt[1], aa[2] * pt[0] - aa[0] * pt[2], aa[0] * pt[1] - aa[1] * pt[0]};
        for (int i = 0; i < 3; ++i) out[i] = pt[i] + wx[i];
    }
}

// rotation.h:72-94
static void aa_to_quaternion(const double* aa, double* q) {
    const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
    if (t2 > std::numeric_limits<double>::epsilon()) {
        const double t = std::sqrt(t2), half = t * 0.5, k = std::sin(half) / t;
        q[0] = std::cos(half);
        q[1] = aa[0] * k;
        q[2] = aa[1] * k;
        q[3] = aa[2] * k;
    } else {
        q[0] = 1.0;
        q[1] = aa[0] * 0.5;
        q[2] = aa[1] * 0.5;
        q[3] = aa[2] * 0.5;
    }
}

// rotation.h:96-131
static void quaternion_to_aa(const double* q, double* aa) {
    const double s2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (s2 > std::numeric_limits<double>::epsilon()) {
        const double st = std::sqrt(s2), ct = q[0];
        const double two_theta = 2.0 * ((ct < 0.0) ? std::atan2(-st, -ct) : std::atan2(st, ct));
        const double k = two_theta / st;
        aa[0] = q[1] * k;
        aa[1] = q[2] * k;
        aa[2] = q[3] * k;
    } else {
        aa[0] = q[1] * 2.0;
        aa[1] = q[2] * 2.0;
        aa[2] = q[3] * 2.0;
    }
}

// BALProblem.cpp:230-246: angle-axis and center c = -R^T t of a camera block.
void host_camera_to_aa_center(const double* cam, bool quat, double aa[3], double center[3]) {
    if (quat)
        quaternion_to_aa(cam, aa);
    else
        for (int i = 0; i < 3; ++i) aa[i] = cam[i];
    const double inv[3] = {-aa[0], -aa[1], -aa[2]};
    const double* t = cam + (quat ? 4 : 3);
    host_rotate_point(inv, t, center);
    for (int i = 0; i < 3; ++i) center[i] *= -1.0;
}

// BALProblem.cpp:248-261: t = -R c.
void host_aa_center_to_camera(const double aa[3], const double center[3], bool quat, double* cam) {
    if (quat)
        aa_to_quaternion(aa, cam);
    else
        for (int i = 0; i < 3; ++i) cam[i] = aa[i];
    double* t = cam + (quat ? 4 : 3);
    host_rotate_point(aa, center, t);
    for (int i = 0; i < 3; ++i) t[i] *= -1.0;
}

namespace {

// BALProblem.cpp:39-55: (double)rand()/RAND_MAX, Marsaglia polar method returning x1*w.
double rand_uniform_glibc() { return static_cast<double>(rand()) / RAND_MAX; }
double rand_normal_glibc() {
    double x1, x2, w;
    do {
        x1 = 2.0 * rand_uniform_glibc() - 1.0;
        x2 = 2.0 * rand_uniform_glibc() - 1.0;
        w = x1 * x1 + x2 * x2;
    } while (w >= 1.0 || w == 0.0);
    w = std::sqrt((-2.0 * std::log(w)) / w);
    return x1 * w;
}

// BALProblem.cpp:71-76: upper median through nth_element at n/2.
double upper_median(std::vector<double>* d) {
    auto mid = d->begin() + d->size() / 2;
    std::nth_element(d->begin(), mid, d->end());
    return *mid;
}

// Whitespace-separated scanner with fscanf("%d"/"%lf") acceptance.
struct Scanner {
    FILE* f;
    bool next_int(int* v) { return fscanf(f, "%d", v) == 1; }
    bool next_double(double* v) { return fscanf(f, "%lf", v) == 1; }
};

}  // namespace

int bal_load(const char* path, bool use_quaternions, ba_bal** out) {
    if (!path || !out) return fail(BA_ERR_INVALID_ARGUMENT, "path/out is NULL");
    FILE* f = fopen(path, "r");
    if (!f) return fail(BA_ERR_IO, "unable to open file %s: %s", path, strerror(errno));
    Scanner sc{f};
    int nc = 0, np = 0, no = 0;
    if (!sc.next_int(&nc) || !sc.next_int(&np) || !sc.next_int(&no) || nc < 0 || np < 0 || no < 0) {
        fclose(f);
        return fail(BA_ERR_IO, "%s: invalid BAL header", path);
    }
    ba_bal* b = new ba_bal();
    b->nc = nc;
    b->np = np;
    b->no = no;
    b->use_quaternions = use_quaternions;
    b->ci.resize(no);
    b->pi.resize(no);
    b->obs.resize(2 * (size_t)no);
    for (int i = 0; i < no; ++i) {
        if (!sc.next_int(&b->ci[i]) || !sc.next_int(&b->pi[i]) || !sc.next_double(&b->obs[2 * (size_t)i]) ||
            !sc.next_double(&b->obs[2 * (size_t)i + 1])) {
            fclose(f);
            delete b;
            return fail(BA_ERR_IO, "%s: truncated observation %d", path, i);
        }
    }
    std::vector<double> params(9 * (size_t)nc + 3 * (size_t)np);
    for (size_t i = 0; i < params.size(); ++i) {
        if (!sc.next_double(&params[i])) {
            fclose(f);
            delete b;
            return fail(BA_ERR_IO, "%s: truncated parameter %zu", path, i);
        }
    }
    fclose(f);
    if (use_quaternions) {  // BALProblem.cpp:116-138
        b->params.resize(10 * (size_t)nc + 3 * (size_t)np);
        for (int c = 0; c < nc; ++c) {
            aa_to_quaternion(&params[9 * (size_t)c], &b->params[10 * (size_t)c]);
            for (int j = 3; j < 9; ++j) b->params[10 * (size_t)c + 1 + j] = params[9 * (size_t)c + j];
        }
        std::copy(params.begin() + 9 * (size_t)nc, params.end(), b->params.begin() + 10 * (size_t)nc);
    } else {
        b->params.swap(params);
    }
    *out = b;
    return BA_OK;
}

int bal_write(const ba_bal* b, const char* path, bool legacy_header) {
    FILE* f = fopen(path, "w");
    if (!f) return fail(BA_ERR_IO, "unable to open file %s for writing: %s", path, strerror(errno));
    if (legacy_header)
        fprintf(f, "%d %d %d %d\n", b->nc, b->nc, b->np, b->no);
    else
        fprintf(f, "%d %d %d\n", b->nc, b->np, b->no);
    for (int i = 0; i < b->no; ++i)
        fprintf(f, "%d %d %g %g\n", b->ci[i], b->pi[i], b->obs[2 * (size_t)i], b->obs[2 * (size_t)i + 1]);
    const int cbs = b->camera_block_size();
    for (int c = 0; c < b->nc; ++c) {
        double cam[9];
        const double* src = &b->params[(size_t)cbs * c];
        if (b->use_quaternions) {
            quaternion_to_aa(src, cam);
            std::memcpy(cam + 3, src + 4, 6 * sizeof(double));
        } else {
            std::memcpy(cam, src, 9 * sizeof(double));
        }
        for (int j = 0; j < 9; ++j) fprintf(f, "%.16g\n", cam[j]);
    }
    const double* pts = b->points();
    for (size_t i = 0; i < 3 * (size_t)b->np; ++i) fprintf(f, "%.16g\n", pts[i]);
    const bool ok = ferror(f) == 0;
    fclose(f);
    return ok ? BA_OK : fail(BA_ERR_IO, "write error on %s", path);
}

int bal_write_ply(const ba_bal* b, const char* path) {
    std::ofstream of(path);
    if (!of) return fail(BA_ERR_IO, "unable to open file %s for writing", path);
    of << "ply" << '\n'
       << "format ascii 1.0" << '\n'
       << "element vertex " << b->nc + b->np << '\n'
       << "property float x" << '\n'
       << "property float y" << '\n'
       << "property float z" << '\n'
       << "property uchar red" << '\n'
       << "property uchar green" << '\n'
       << "property uchar blue" << '\n'
       << "end_header" << std::endl;
    const int cbs = b->camera_block_size();
    for (int c = 0; c < b->nc; ++c) {  // camera centers in green
        double aa[3], center[3];
        host_camera_to_aa_center(&b->params[(size_t)cbs * c], b->use_quaternions, aa, center);
        of << center[0] << ' ' << center[1] << ' ' << center[2] << " 0 255 0" << '\n';
    }
    const double* pts = b->points();
    for (int i = 0; i < b->np; ++i) {  // structure in white
        for (int j = 0; j < 3; ++j) of << pts[3 * (size_t)i + j] << ' ';
        of << " 255 255 255\n";
    }
    of.close();
    return of ? BA_OK : fail(BA_ERR_IO, "write error on %s", path);
}

void bal_normalize(ba_bal* b) {
    const int np = b->np, nc = b->nc;
    if (np == 0) return;
    double* pts = b->points();
    std::vector<double> tmp(np);
    double med[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < np; ++j) tmp[j] = pts[3 * (size_t)j + i];
        med[i] = upper_median(&tmp);
    }
    for (int j = 0; j < np; ++j) {
        const double* p = pts + 3 * (size_t)j;
        tmp[j] = std::fabs(p[0] - med[0]) + std::fabs(p[1] - med[1]) + std::fabs(p[2] - med[2]);
    }
    const double scale = 100.0 / upper_median(&tmp);
    for (int j = 0; j < np; ++j)
        for (int i = 0; i < 3; ++i) pts[3 * (size_t)j + i] = scale * (pts[3 * (size_t)j + i] - med[i]);
    const int cbs = b->camera_block_size();
    for (int c = 0; c < nc; ++c) {
        double* cam = &b->params[(size_t)cbs * c];
        double aa[3], center[3];
        host_camera_to_aa_center(cam, b->use_quaternions, aa, center);
        for (int i = 0; i < 3; ++i) center[i] = scale * (center[i] - med[i]);
        host_aa_center_to_camera(aa, center, b->use_quaternions, cam);
    }
}

void bal_perturb(ba_bal* b, double rs, double ts, double ps) {
    double* pts = b->points();
    if (ps > 0)
        for (int i = 0; i < b->np; ++i)
            for (int j = 0; j < 3; ++j) pts[3 * (size_t)i + j] += rand_normal_glibc() * ps;
    const int cbs = b->camera_block_size();
    for (int c = 0; c < b->nc; ++c) {
        double* cam = &b->params[(size_t)cbs * c];
        double aa[3], center[3];
        host_camera_to_aa_center(cam, b->use_quaternions, aa, center);
        if (rs > 0.0)
            for (int j = 0; j < 3; ++j) aa[j] += rand_normal_glibc() * rs;
        host_aa_center_to_camera(aa, center, b->use_quaternions, cam);
        if (ts > 0.0)
            for (int j = 0; j < 3; ++j) cam[cbs - 6 + j] += rand_normal_glibc() * ts;
    }
}

}  // namespace mba

using namespace mba;

extern "C" {

int ba_bal_load(const char* path, int32_t use_quaternions, ba_bal** out) {
    clear_error();
    return bal_load(path, use_quaternions != 0, out);
}

int ba_bal_from_arrays(int32_t nc, int32_t np, int32_t no, const int32_t* ci, const int32_t* pi, const double* obs,
                       const double* cams, const double* pts, ba_bal** out) {
    clear_error();
    if (!out || nc < 0 || np < 0 || no < 0) return fail(BA_ERR_INVALID_ARGUMENT, "invalid arguments");
    if ((no > 0 && (!ci || !pi || !obs)) || (nc > 0 && !cams) || (np > 0 && !pts))
        return fail(BA_ERR_INVALID_ARGUMENT, "NULL array");
    ba_bal* b = new ba_bal();
    b->nc = nc;
    b->np = np;
    b->no = no;
    b->ci.assign(ci, ci + no);
    b->pi.assign(pi, pi + no);
    b->obs.assign(obs, obs + 2 * (size_t)no);
    b->params.resize(9 * (size_t)nc + 3 * (size_t)np);
    std::copy(cams, cams + 9 * (size_t)nc, b->params.begin());
    std::copy(pts, pts + 3 * (size_t)np, b->params.begin() + 9 * (size_t)nc);
    *out = b;
    return BA_OK;
}

int ba_bal_view(ba_bal* b, ba_problem_view* v) {
    clear_error();
    if (!b || !v) return fail(BA_ERR_INVALID_ARGUMENT, "NULL handle/view");
    *v = b->view();
    return BA_OK;
}

int ba_bal_normalize(ba_bal* b) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL handle");
    bal_normalize(b);
    return BA_OK;
}

int ba_bal_perturb(ba_bal* b, double rs, double ts, double ps) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!(rs >= 0.0) || !(ts >= 0.0) || !(ps >= 0.0)) return fail(BA_ERR_INVALID_ARGUMENT, "negative sigma");
    bal_perturb(b, rs, ts, ps);
    return BA_OK;
}

int ba_bal_write(ba_bal* b, const char* path, int32_t legacy_header) {
    clear_error();
    if (!b || !path) return fail(BA_ERR_INVALID_ARGUMENT, "NULL handle/path");
    return bal_write(b, path, legacy_header != 0);
}

int ba_bal_write_ply(ba_bal* b, const char* path) {
    clear_error();
    if (!b || !path) return fail(BA_ERR_INVALID_ARGUMENT, "NULL handle/path");
    return bal_write_ply(b, path);
}

int ba_bal_destroy(ba_bal* b) {
    delete b;
    return BA_OK;
}

}  // extern "C"

// ---- slamplay_amd::BALProblem ----
namespace slamplay_amd {

BALProblem::BALProblem(const std::string& filename, bool use_quaternions) : use_quaternions_(use_quaternions) {
    if (ba_bal_load(filename.c_str(), use_quaternions ? 1 : 0, &handle_) != BA_OK)
        throw std::runtime_error(ba_last_error_string());
    view_ = handle_->view();
}

BALProblem::BALProblem(ba_bal* handle) : handle_(handle) {
    if (!handle_) throw std::invalid_argument("BALProblem: NULL handle");
    use_quaternions_ = handle_->use_quaternions;
    view_ = handle_->view();
}

BALProblem::~BALProblem() { ba_bal_destroy(handle_); }

void BALProblem::WriteToFile(const std::string& filename) const {
    if (mba::bal_write(handle_, filename.c_str(), false) != BA_OK) throw std::runtime_error(ba_last_error_string());
}

void BALProblem::WriteToPLYFile(const std::string& filename) const {
    if (mba::bal_write_ply(handle_, filename.c_str()) != BA_OK) throw std::runtime_error(ba_last_error_string());
}

void BALProblem::Normalize() { mba::bal_normalize(handle_); }

void BALProblem::Perturb(double rotation_sigma, double translation_sigma, double point_sigma) {
    mba::bal_perturb(handle_, rotation_sigma, translation_sigma, point_sigma);
}

}  // namespace slamplay_amd
