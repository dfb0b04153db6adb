// synth.cpp -- seeded synthetic BA problems for the benchmark configurations.
//
// Camera model and noise (all configurations): Snavely camera (angle-axis, t, f, k1, k2) looking
// down -z (utils/camera/SnavelyReprojectionError.h:72-73), f ~ N(focal_mean, focal_sigma^2),
// k1, k2 ~ N(0, distortion_sigma^2), observation = exact projection + N(0, pixel_noise^2) per axis.
// Initial guess: points += N(0, point_sigma^2); per camera the angle-axis is perturbed with the
// camera center held fixed, then t += N(0, translation_sigma^2) -- the order of
// BALProblem::Perturb (utils/datasets/BALProblem.cpp:305-336) with this generator's own RNG.
//
// kind 1 (circle): cameras on a radius-10 circle in z = 0 facing the origin, points ~ N(0, 2^2 I).
//   A camera observes a point when it lies in front of it, within 30 degrees of the optical axis
//   (a 60-degree field of view) and on the camera's side of the origin (X . c_hat > 0).
// kind 2 (random walk): unit steps with heading noise N(0, 0.05^2) rad, position noise
//   N(0, 0.05^2), orientation noise N(0, 0.02^2) rad per axis, cameras facing the heading.  Each
//   point has a track of t consecutive cameras, t ~ Zipf(zipf_s) truncated and renormalised on
//   [min_track, max_track] (uniform when zipf_s == 0), placed 5-25 units ahead of its last camera
//   with N(0, 3^2) lateral offsets and kept only where every observing camera sees it in front
//   within 45 degrees of its axis.  With num_points == 0 tracks are drawn until exactly
//   target_observations observations exist.
// Points with fewer than two observations are dropped; observations are ordered by point, then
// camera, as in BAL files.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bal_host.h"
#include "common.h"

namespace mba {
namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    bool has_spare = false;
    double spare = 0.0;
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int uniform_int(int lo, int hi) { return lo + (int)(uniform() * (double)(hi - lo + 1)); }  // [lo, hi]
    double normal() {  // Box-Muller
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        spare = rad * std::sin(2.0 * M_PI * u2);
        has_spare = true;
        return rad * std::cos(2.0 * M_PI * u2);
    }
};

struct Mat3 {
    double m[3][3];
};

void normalize3(double v[3]) {
    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] /= n;
}
void cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// World-to-camera rotation whose -z axis looks along `forward`.
Mat3 look_along(const double forward[3]) {
    double rz[3] = {-forward[0], -forward[1], -forward[2]};
    normalize3(rz);
    const double up[3] = {0.0, 0.0, 1.0};
    double rx[3], ry[3];
    cross(up, rz, rx);
    normalize3(rx);
    cross(rz, rx, ry);
    Mat3 R;
    for (int j = 0; j < 3; ++j) {
        R.m[0][j] = rx[j];
        R.m[1][j] = ry[j];
        R.m[2][j] = rz[j];
    }
    return R;
}

Mat3 mul(const Mat3& A, const Mat3& B) {
    Mat3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return C;
}

Mat3 exp_so3(const double w[3]) {
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    Mat3 R{};
    for (int i = 0; i < 3; ++i) R.m[i][i] = 1.0;
    if (th == 0.0) return R;
    const double k[3] = {w[0] / th, w[1] / th, w[2] / th};
    const double K[3][3] = {{0, -k[2], k[1]}, {k[2], 0, -k[0]}, {-k[1], k[0], 0}};
    const double s = std::sin(th), c1 = 1.0 - std::cos(th);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double kk = 0.0;
            for (int l = 0; l < 3; ++l) kk += K[i][l] * K[l][j];
            R.m[i][j] += s * K[i][j] + c1 * kk;
        }
    return R;
}

// Rotation matrix -> angle-axis through a unit quaternion (largest-pivot extraction), so that
// rotation angles near pi stay accurate.
void log_so3(const Mat3& Rm, double aa[3]) {
    const double(*R)[3] = Rm.m;
    const double tr = R[0][0] + R[1][1] + R[2][2];
    double q[4];
    if (tr > R[0][0] && tr > R[1][1] && tr > R[2][2]) {
        const double w = 0.5 * std::sqrt(1.0 + tr);
        q[0] = w;
        q[1] = (R[2][1] - R[1][2]) / (4.0 * w);
        q[2] = (R[0][2] - R[2][0]) / (4.0 * w);
        q[3] = (R[1][0] - R[0][1]) / (4.0 * w);
    } else if (R[0][0] >= R[1][1] && R[0][0] >= R[2][2]) {
        const double x = 0.5 * std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]);
        q[1] = x;
        q[0] = (R[2][1] - R[1][2]) / (4.0 * x);
        q[2] = (R[0][1] + R[1][0]) / (4.0 * x);
        q[3] = (R[0][2] + R[2][0]) / (4.0 * x);
    } else if (R[1][1] >= R[2][2]) {
        const double y = 0.5 * std::sqrt(1.0 - R[0][0] + R[1][1] - R[2][2]);
        q[2] = y;
        q[0] = (R[0][2] - R[2][0]) / (4.0 * y);
        q[1] = (R[0][1] + R[1][0]) / (4.0 * y);
        q[3] = (R[1][2] + R[2][1]) / (4.0 * y);
    } else {
        const double z = 0.5 * std::sqrt(1.0 - R[0][0] - R[1][1] + R[2][2]);
        q[3] = z;
        q[0] = (R[1][0] - R[0][1]) / (4.0 * z);
        q[1] = (R[0][2] + R[2][0]) / (4.0 * z);
        q[2] = (R[1][2] + R[2][1]) / (4.0 * z);
    }
    if (q[0] < 0)
        for (double& v : q) v = -v;
    const double sn = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (sn < 1e-300) {
        aa[0] = aa[1] = aa[2] = 0.0;
        return;
    }
    const double angle = 2.0 * std::atan2(sn, q[0]);
    for (int i = 0; i < 3; ++i) aa[i] = q[1 + i] / sn * angle;
}

struct Cam {
    double p[9];   // angle-axis, t, f, k1, k2
    double C[3];   // center
    Mat3 R;
};

void make_camera(Rng& rng, const ba_synth_params& s, const Mat3& R, const double C[3], Cam* cam) {
    cam->R = R;
    for (int i = 0; i < 3; ++i) cam->C[i] = C[i];
    log_so3(R, cam->p);
    double rc[3];
    host_rotate_point(cam->p, C, rc);
    for (int i = 0; i < 3; ++i) cam->p[3 + i] = -rc[i];
    cam->p[6] = s.focal_mean + s.focal_sigma * rng.normal();
    cam->p[7] = s.distortion_sigma * rng.normal();
    cam->p[8] = s.distortion_sigma * rng.normal();
}

// Camera-frame point p = R(w) X + t using the reference rotation.
void to_camera(const Cam& c, const double X[3], double p[3]) {
    host_rotate_point(c.p, X, p);
    for (int i = 0; i < 3; ++i) p[i] += c.p[3 + i];
}

void project(const Cam& c, const double X[3], double uv[2]) {
    double p[3];
    to_camera(c, X, p);
    const double xp = -p[0] / p[2], yp = -p[1] / p[2];
    const double r2 = xp * xp + yp * yp;
    const double d = 1.0 + r2 * (c.p[7] + c.p[8] * r2);
    uv[0] = c.p[6] * d * xp;
    uv[1] = c.p[6] * d * yp;
}

struct Track {
    double X[3];
    std::vector<int> cams;
};

}  // namespace

int synth_generate(const ba_synth_params& s, ba_bal** out, double* gt, int64_t gt_cap) {
    if (s.num_cameras < 2) return fail(BA_ERR_INVALID_ARGUMENT, "synthetic problem needs >= 2 cameras");
    if (s.kind != 1 && s.kind != 2) return fail(BA_ERR_INVALID_ARGUMENT, "unknown synthetic kind %d", s.kind);
    Rng rng(s.seed);
    const int nc = s.num_cameras;
    std::vector<Cam> cams(nc);
    std::vector<Track> tracks;
    if (s.kind == 1) {
        for (int i = 0; i < nc; ++i) {
            const double phi = 2.0 * M_PI * i / nc;
            const double C[3] = {10.0 * std::cos(phi), 10.0 * std::sin(phi), 0.0};
            const double fwd[3] = {-C[0], -C[1], -C[2]};
            make_camera(rng, s, look_along(fwd), C, &cams[i]);
        }
        const double tan_half = std::tan(30.0 * M_PI / 180.0);
        for (int j = 0; j < s.num_points; ++j) {
            Track t;
            for (int k = 0; k < 3; ++k) t.X[k] = 2.0 * rng.normal();
            for (int i = 0; i < nc; ++i) {
                double p[3];
                to_camera(cams[i], t.X, p);
                const double side = t.X[0] * cams[i].C[0] + t.X[1] * cams[i].C[1] + t.X[2] * cams[i].C[2];
                if (p[2] < 0.0 && std::hypot(p[0], p[1]) <= tan_half * (-p[2]) && side > 0.0) t.cams.push_back(i);
            }
            if (t.cams.size() >= 2) tracks.push_back(std::move(t));
        }
    } else {
        double psi = 0.0, C[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < nc; ++i) {
            if (i > 0) {
                psi += 0.05 * rng.normal();
                C[0] += std::cos(psi) + 0.05 * rng.normal();
                C[1] += std::sin(psi) + 0.05 * rng.normal();
                C[2] += 0.05 * rng.normal();
            }
            const double fwd[3] = {std::cos(psi), std::sin(psi), 0.0};
            const double wn[3] = {0.02 * rng.normal(), 0.02 * rng.normal(), 0.02 * rng.normal()};
            make_camera(rng, s, mul(exp_so3(wn), look_along(fwd)), C, &cams[i]);
        }
        const int tmin = std::max(2, s.min_track), tmax = std::max(tmin, std::min(s.max_track, nc));
        std::vector<double> cdf;
        for (int t = tmin; t <= tmax; ++t) cdf.push_back((cdf.empty() ? 0.0 : cdf.back()) + std::pow((double)t, -s.zipf_s));
        for (double& v : cdf) v /= cdf.back();
        const bool by_obs = s.num_points <= 0;
        if (by_obs && s.target_observations < 2) return fail(BA_ERR_INVALID_ARGUMENT, "target_observations < 2");
        int64_t nobs = 0;
        int64_t guard = 0;
        while (by_obs ? nobs < s.target_observations : (int)tracks.size() < s.num_points) {
            if (++guard > 100LL * std::max<int64_t>(1, by_obs ? s.target_observations : s.num_points))
                return fail(BA_ERR_INVALID_ARGUMENT, "synthetic generator could not place points");
            int len = tmin + (int)(std::lower_bound(cdf.begin(), cdf.end(), rng.uniform()) - cdf.begin());
            len = std::min(len, tmax);
            if (by_obs) {
                const int64_t left = s.target_observations - nobs;
                if (left < len) len = (int)left;
                if (len < 2) len = 2;
            }
            const int a = rng.uniform_int(0, nc - len);
            const Cam& last = cams[a + len - 1];
            Track t;
            bool ok = false;
            for (int attempt = 0; attempt < 10 && !ok; ++attempt) {
                const double d = rng.uniform(5.0, 25.0), ox = 3.0 * rng.normal(), oy = 3.0 * rng.normal();
                for (int k = 0; k < 3; ++k)
                    t.X[k] = last.C[k] - d * last.R.m[2][k] + ox * last.R.m[0][k] + oy * last.R.m[1][k];
                ok = true;
                for (int i = a; i < a + len && ok; ++i) {
                    double p[3];
                    to_camera(cams[i], t.X, p);
                    ok = p[2] < -0.5 && std::fabs(p[0]) <= -p[2] && std::fabs(p[1]) <= -p[2];
                }
            }
            if (!ok) continue;
            for (int i = a; i < a + len; ++i) t.cams.push_back(i);
            nobs += len;
            tracks.push_back(std::move(t));
        }
    }
    const int np = (int)tracks.size();
    int64_t no64 = 0;
    for (const Track& t : tracks) no64 += (int64_t)t.cams.size();
    if (no64 > INT32_MAX) return fail(BA_ERR_CAPACITY, "too many observations");
    const int no = (int)no64;
    ba_bal* b = new ba_bal();
    b->nc = nc;
    b->np = np;
    b->no = no;
    b->ci.resize(no);
    b->pi.resize(no);
    b->obs.resize(2 * (size_t)no);
    b->params.resize(9 * (size_t)nc + 3 * (size_t)np);
    for (int i = 0; i < nc; ++i) std::memcpy(&b->params[9 * (size_t)i], cams[i].p, 9 * sizeof(double));
    int k = 0;
    for (int j = 0; j < np; ++j) {
        std::memcpy(&b->params[9 * (size_t)nc + 3 * (size_t)j], tracks[j].X, 3 * sizeof(double));
        for (int c : tracks[j].cams) {
            double uv[2];
            project(cams[c], tracks[j].X, uv);
            b->ci[k] = c;
            b->pi[k] = j;
            b->obs[2 * (size_t)k] = uv[0] + s.pixel_noise * rng.normal();
            b->obs[2 * (size_t)k + 1] = uv[1] + s.pixel_noise * rng.normal();
            ++k;
        }
    }
    if (gt) {
        if (gt_cap < (int64_t)b->params.size()) {
            delete b;
            return fail(BA_ERR_CAPACITY, "ground truth buffer holds %lld doubles, need %zu", (long long)gt_cap,
                        b->params.size());
        }
        std::memcpy(gt, b->params.data(), b->params.size() * sizeof(double));
    }
    // initial guess: points, then per camera rotation about the fixed center, then translation
    double* pts = b->points();
    for (size_t i = 0; i < 3 * (size_t)np; ++i) pts[i] += s.point_sigma * rng.normal();
    for (int i = 0; i < nc; ++i) {
        double* cam = &b->params[9 * (size_t)i];
        double aa[3], center[3];
        host_camera_to_aa_center(cam, false, aa, center);
        for (int j = 0; j < 3; ++j) aa[j] += s.rotation_sigma * rng.normal();
        host_aa_center_to_camera(aa, center, false, cam);
        for (int j = 0; j < 3; ++j) cam[3 + j] += s.translation_sigma * rng.normal();
    }
    *out = b;
    return BA_OK;
}

}  // namespace mba

extern "C" {

int ba_synth_default(int32_t config, ba_synth_params* p) {
    mba::clear_error();
    if (!p) return mba::fail(BA_ERR_INVALID_ARGUMENT, "params is NULL");
    *p = ba_synth_params{};
    p->pixel_noise = 0.5;
    p->rotation_sigma = 0.01;
    p->translation_sigma = 0.1;
    p->point_sigma = 0.1;
    p->focal_mean = 500.0;
    p->focal_sigma = 20.0;
    p->distortion_sigma = 1e-3;
    p->seed = 1;
    switch (config) {
        case 1:
            p->kind = 1;
            p->num_cameras = 16;
            p->num_points = 2000;
            break;
        case 2:
            p->kind = 2;
            p->num_cameras = 250;
            p->num_points = 60000;
            p->min_track = 2;
            p->max_track = 40;
            p->zipf_s = 1.5;
            break;
        case 3:
            p->kind = 2;
            p->num_cameras = 1500;
            p->num_points = 0;
            p->target_observations = 1200000;
            p->min_track = 2;
            p->max_track = 60;
            p->zipf_s = 1.3;
            break;
        case 4:
        case 5:
            p->kind = 2;
            p->num_cameras = 12;
            p->num_points = 1000;
            p->min_track = 2;
            p->max_track = 12;
            p->zipf_s = 0.0;
            break;
        default:
            return mba::fail(BA_ERR_INVALID_ARGUMENT, "unknown configuration %d", config);
    }
    return BA_OK;
}

int ba_synth_generate(const ba_synth_params* p, ba_bal** out, double* gt, int64_t gt_cap) {
    mba::clear_error();
    if (!p || !out) return mba::fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    return mba::synth_generate(*p, out, gt, gt_cap);
}

}  // extern "C"
