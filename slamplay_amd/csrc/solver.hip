// solver.hip -- host side of the device-resident LM bundle-adjustment solve (ba_solve and the
// ba_solver_* handle API).  Replaces SolveBA(BALProblem&) of
// backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp:58-101: one Snavely residual
// per observation (:70-87) with HuberLoss(1.0) (:78), parameters updated in place (:83-84),
// Levenberg-Marquardt with points eliminated by the Schur complement (:96-100).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include <hip/hip_ext.h>

#include "dense.h"
#include "lm_kernels.h"
#include "preprocess.h"
#include "solver.h"
#include "stereo.h"

namespace mba {

// the instantiation of kernel k for this solver's residual model (d.model)
#define BA_MODEL_KERNEL(k) pick_model(d.model, k<Real, kModelSnavely>, k<Real, kModelSnavelyG2o>, k<Real, kModelStereo>)

namespace {

Opts make_opts(const ba_options& o) {
    Opts r{};
    r.g2o = o.mode == BA_MODE_G2O;
    r.g2o_max_trials = o.g2o_max_trials;
    r.g2o_tau = o.g2o_tau;
    r.loss = o.loss;
    r.max_iterations = o.max_num_iterations;
    r.max_invalid = o.max_num_consecutive_invalid_steps;
    r.jacobi_scaling = o.jacobi_scaling;
    r.pcg_max_iterations = o.pcg_max_iterations;
    r.loss_a = o.loss_scale;
    r.ftol = o.function_tolerance;
    r.gtol = o.gradient_tolerance;
    r.ptol = o.parameter_tolerance;
    r.max_radius = o.max_trust_region_radius;
    r.min_radius = o.min_trust_region_radius;
    r.min_rel_decrease = o.min_relative_decrease;
    r.min_diag = o.min_lm_diagonal;
    r.max_diag = o.max_lm_diagonal;
    r.pcg_tol = o.pcg_tolerance;
    if (r.g2o) {
        // g2o damps H + lambda I without column scaling: D^2 = clamp(diag, 1, 1) / radius = lambda
        r.jacobi_scaling = 0;
        r.min_diag = r.max_diag = 1.0;
    }
    return r;
}

int check_options(const ba_options* o) {
    if (!o) return fail(BA_ERR_INVALID_ARGUMENT, "options is NULL");
    if (o->max_num_iterations < 0) return fail(BA_ERR_INVALID_ARGUMENT, "max_num_iterations < 0");
    if (o->max_num_consecutive_invalid_steps < 1) return fail(BA_ERR_INVALID_ARGUMENT, "max_num_consecutive_invalid_steps < 1");
    if (o->loss != BA_LOSS_NONE && o->loss != BA_LOSS_HUBER) return fail(BA_ERR_INVALID_ARGUMENT, "unknown loss %d", o->loss);
    if (o->loss == BA_LOSS_HUBER && !(o->loss_scale > 0.0)) return fail(BA_ERR_INVALID_ARGUMENT, "loss_scale must be > 0");
    if (o->precision != BA_PRECISION_FP64 && o->precision != BA_PRECISION_MIXED)
        return fail(BA_ERR_INVALID_ARGUMENT, "unknown precision %d", o->precision);
    if (o->linear_solver < BA_SOLVER_AUTO || o->linear_solver > BA_SOLVER_ITERATIVE_SCHUR)
        return fail(BA_ERR_INVALID_ARGUMENT, "unknown linear solver %d", o->linear_solver);
    if (!(o->initial_trust_region_radius > 0.0)) return fail(BA_ERR_INVALID_ARGUMENT, "initial_trust_region_radius must be > 0");
    if (!(o->min_lm_diagonal > 0.0) || !(o->max_lm_diagonal >= o->min_lm_diagonal))
        return fail(BA_ERR_INVALID_ARGUMENT, "invalid LM diagonal bounds");
    if (o->pcg_max_iterations < 0) return fail(BA_ERR_INVALID_ARGUMENT, "pcg_max_iterations < 0");
    if (o->launch_mode != 0 && o->launch_mode != 1) return fail(BA_ERR_INVALID_ARGUMENT, "unknown launch_mode %d", o->launch_mode);
    if (o->mode != BA_MODE_CERES && o->mode != BA_MODE_G2O) return fail(BA_ERR_INVALID_ARGUMENT, "unknown mode %d", o->mode);
    if (o->mode == BA_MODE_G2O && (o->g2o_max_trials < 1 || !(o->g2o_tau > 0.0)))
        return fail(BA_ERR_INVALID_ARGUMENT, "g2o mode needs g2o_max_trials >= 1 and g2o_tau > 0");
    return BA_OK;
}

}  // namespace

int validate_options(const ba_options* o) { return check_options(o); }

Opts solver_opts(const ba_options& o) { return make_opts(o); }

// ---------------------------------------------------------------------------------------------
// Host-built device layout shared by the FP64 and mixed solvers.
int build_layout(const ba_problem_view* v, Layout* L) {
    const int nc = v->num_cameras, np = v->num_points, no = v->num_observations;
    L->nc = nc;
    L->np = np;
    L->no = no;
    std::vector<int> track(np, 0), first(np, nc);
    for (int k = 0; k < no; ++k) {
        track[v->point_index[k]]++;
        first[v->point_index[k]] = std::min(first[v->point_index[k]], v->camera_index[k]);
    }
    // points by first observing camera (a camera's observations then touch clustered points),
    // then, inside windows of 512 points, by track length (longest first) so that a wave's 64
    // points have similar tracks and the SELL padding stays small
    L->pt_perm.resize(np);
    std::iota(L->pt_perm.begin(), L->pt_perm.end(), 0);
    std::sort(L->pt_perm.begin(), L->pt_perm.end(), [&](int a, int b) {
        if (first[a] != first[b]) return first[a] < first[b];
        if (track[a] != track[b]) return track[a] > track[b];
        return a < b;
    });
    constexpr int kWindow = 512;
    for (int s0 = 0; s0 < np; s0 += kWindow)
        std::stable_sort(L->pt_perm.begin() + s0, L->pt_perm.begin() + std::min(np, s0 + kWindow),
                         [&](int a, int b) { return track[a] > track[b]; });
    std::vector<int> inv(np);
    for (int i = 0; i < np; ++i) inv[L->pt_perm[i]] = i;
    // observations by (camera, internal point)
    std::vector<int64_t> key(no);
    for (int k = 0; k < no; ++k) key[k] = ((int64_t)v->camera_index[k] << 32) | (int64_t)inv[v->point_index[k]];
    L->co_orig.resize(no);
    std::iota(L->co_orig.begin(), L->co_orig.end(), 0);
    std::sort(L->co_orig.begin(), L->co_orig.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
    L->cam_off.assign(nc + 1, 0);
    L->co_pt.resize(no);
    L->co_uv.resize(2 * (size_t)no);
    for (int j = 0; j < no; ++j) {
        const int k = L->co_orig[j];
        L->cam_off[v->camera_index[k] + 1]++;
        L->co_pt[j] = inv[v->point_index[k]];
        L->co_uv[2 * (size_t)j] = v->observations[2 * (size_t)k];
        L->co_uv[2 * (size_t)j + 1] = v->observations[2 * (size_t)k + 1];
    }
    for (int c = 0; c < nc; ++c) L->cam_off[c + 1] += L->cam_off[c];
    // SELL-64 slots of camera-order indices per internal point
    L->n_chunks = (np + 63) / 64;
    L->chunk_len.assign(L->n_chunks, 0);
    L->chunk_off.assign(L->n_chunks + 1, 0);
    for (int ch = 0; ch < L->n_chunks; ++ch) {
        int len = 0;
        for (int i = ch * 64; i < std::min(np, ch * 64 + 64); ++i) len = std::max(len, track[L->pt_perm[i]]);
        L->chunk_len[ch] = len;
        L->chunk_off[ch + 1] = L->chunk_off[ch] + 64 * L->chunk_len[ch];
    }
    L->slot_j.assign((size_t)L->chunk_off[L->n_chunks], -1);
    L->slot_of.assign(no, -1);
    std::vector<int> fill(np, 0);
    for (int j = 0; j < no; ++j) {
        const int ip = L->co_pt[j];
        const int ch = ip >> 6, lane = ip & 63;
        const size_t slot = (size_t)L->chunk_off[ch] + 64 * (size_t)fill[ip] + lane;
        L->slot_j[slot] = j;
        L->slot_of[j] = (int)slot;
        fill[ip]++;
    }
    return BA_OK;
}

void build_pair_blocks(const Layout& L, PairBlocks* out) {
    const int nc = L.nc;
    std::vector<int> cam_of(L.no);
    for (int c = 0; c < nc; ++c)
        for (int j = L.cam_off[c]; j < L.cam_off[c + 1]; ++j) cam_of[j] = c;
    std::vector<int> pt_off(L.np + 1, 0);
    for (int j = 0; j < L.no; ++j) pt_off[L.co_pt[j] + 1]++;
    for (int p = 0; p < L.np; ++p) pt_off[p + 1] += pt_off[p];
    std::vector<int> pt_obs(L.no), fill(pt_off.begin(), pt_off.end() - 1);
    for (int j = 0; j < L.no; ++j) pt_obs[fill[L.co_pt[j]]++] = j;
    // counting sort of the pairs (i, j), camera(i) <= camera(j), of every point by block key
    // camera(i) * nc + camera(j); within a block the pairs keep point order
    const size_t nkey = (size_t)nc * nc;
    std::vector<int64_t> cnt(nkey + 1, 0);
    for (int p = 0; p < L.np; ++p)
        for (int x = pt_off[p]; x < pt_off[p + 1]; ++x)
            for (int y = pt_off[p]; y < pt_off[p + 1]; ++y) {
                const int a = cam_of[pt_obs[x]], b = cam_of[pt_obs[y]];
                if (a <= b) cnt[(size_t)a * nc + b + 1]++;
            }
    for (size_t k = 0; k < nkey; ++k) cnt[k + 1] += cnt[k];
    PairBlocks& o = *out;
    o = PairBlocks{};
    o.i.resize(cnt[nkey]);
    o.j.resize(cnt[nkey]);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int p = 0; p < L.np; ++p)
            for (int x = pt_off[p]; x < pt_off[p + 1]; ++x)
                for (int y = pt_off[p]; y < pt_off[p + 1]; ++y) {
                    const int i = pt_obs[x], j = pt_obs[y];
                    const int a = cam_of[i], b = cam_of[j];
                    if (a <= b) {
                        const int64_t q = pos[(size_t)a * nc + b]++;
                        o.i[q] = i;
                        o.j[q] = j;
                    }
                }
    }
    // blocks (a, b) with pairs, in key order; every diagonal block is listed
    for (int a = 0; a < nc; ++a)
        for (int b = a; b < nc; ++b) {
            const size_t k = (size_t)a * nc + b;
            if (a == b || cnt[k + 1] > cnt[k]) {
                o.a.push_back(a);
                o.b.push_back(b);
                o.off.push_back((int)cnt[k]);
            }
        }
    o.off.push_back((int)cnt[nkey]);
}

// ---------------------------------------------------------------------------------------------
// Threads per workgroup of k_jz_dir and k_jtx_dir (one workgroup per camera).  128 for k_jz_dir:
// at its register use a CU then holds eight of them, so the 1,500 cameras of the 1.2M-observation
// problem run as one wave of workgroups instead of two (16.1 against 17.4 us per launch measured).
constexpr int kJzWG = 128;
constexpr int kJtxWG = 256;

template <typename Real>
struct Solver final : SolverBase {
    ba_options opt{};
    hipStream_t s = nullptr;
    int device = 0;
    int linear_solver = BA_SOLVER_ITERATIVE_SCHUR;
    double clock_khz = 100000.0;  // rate of the device wall clock (trace row times)
    Layout L;
    Dev<Real> d{};
    DevBuf<int> cam_off, co_pt, chunk_off, chunk_len, slot_of, co_side;
    DevBuf<double> co_uv64;
    DevBuf<Real> co_uvr;
    DevBuf<R4<Real>> co_rec;  // per observation: X - c and the Huber weight (direction-form PCG)
    DevBuf<double> cam_x, pt_x, cam_x0, pt_x0, cam_c, pt_c, dcam, dpt;
    DevBuf<CameraPrep<Real>> prep;
    DevBuf<CameraPrep<double>> prep_c;
    DevBuf<double> U, gc, sc, dc2, M, Minv, bc, px, pr, pz, pp, qc;
    DevBuf<double> V, gp, sp, dp2, Vinv64, upd_part, rz_hist;
    DevBuf<Real> VinvR;
    DevBuf<R4<Real>> qp, u;
    DevBuf<double4> zs;
    DevBuf<double4> pv;
    DevBuf<double> cam_part, pt_part;
    DevBuf<LMState> st;
    DevBuf<ba_iteration> trace;
    DevBuf<int> pt_orig;  // internal point -> caller point, uploaded on the first Normalize / Perturb
    PrepScratch prep_w;
    LMState* st_host = nullptr;  // pinned
    double wait_timeout_s = 0.0;
    hipEvent_t ev_wait = nullptr;
    DenseSchur<Real> dense;
    // one LM iteration captured once and replayed (launch_mode 0)
    hipGraphExec_t iter_graph = nullptr;
    // profiling
    bool profiling = false;
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<ba_kernel_time> ktimes;

    ~Solver() override {
        if (iter_graph) (void)hipGraphExecDestroy(iter_graph);
        if (st_host) (void)hipHostFree(st_host);
        if (ev_wait) (void)hipEventDestroy(ev_wait);
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }

    int64_t device_bytes() const override {
        int64_t b = 0;
        b += cam_off.bytes() + co_pt.bytes() + chunk_off.bytes() + chunk_len.bytes() + slot_of.bytes();
        b += co_uv64.bytes() + co_uvr.bytes() + co_rec.bytes();
        b += cam_x.bytes() * 4 + pt_x.bytes() * 4 + dcam.bytes() + dpt.bytes();
        b += prep.bytes() + prep_c.bytes();
        b += M.bytes() * 2 + gc.bytes() * 10 + upd_part.bytes() + rz_hist.bytes() + dp2.bytes();
        b += V.bytes() + gp.bytes() + sp.bytes() + Vinv64.bytes();
        b += VinvR.bytes() + qp.bytes() + zs.bytes() + pv.bytes() + u.bytes();
        b += cam_part.bytes() + pt_part.bytes() + dense.device_bytes();
        return b;
    }
    int solver_kind() const override { return linear_solver; }
    void set_wait_timeout(double seconds) override { wait_timeout_s = seconds; }
    int sync(const char* what) {
        if (wait_timeout_s > 0.0 && !ev_wait) BA_HIP(hipEventCreateWithFlags(&ev_wait, hipEventDisableTiming));
        return wait_stream(s, ev_wait, wait_timeout_s, what);
    }

    int create(const ba_problem_view* v, const ba_options& o, const StereoInput* stereo = nullptr) {
        opt = o;
        device = o.device;
        BA_TRY(use_device(device));
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            clock_khz = khz;
        BA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        BA_HIP(hipHostMalloc((void**)&st_host, sizeof(LMState), hipHostMallocDefault));
        BA_TRY(build_layout(v, &L));
        const int nc = L.nc, np = L.np, no = L.no;
        linear_solver = o.linear_solver;
        if (linear_solver == BA_SOLVER_AUTO)
            linear_solver = nc <= o.dense_schur_max_cameras ? BA_SOLVER_DENSE_SCHUR : BA_SOLVER_ITERATIVE_SCHUR;
        BA_TRY(cam_off.upload(L.cam_off.data(), L.cam_off.size(), s));
        BA_TRY(co_pt.upload(L.co_pt.data(), L.co_pt.size(), s));
        BA_TRY(chunk_off.upload(L.chunk_off.data(), L.chunk_off.size(), s));
        BA_TRY(chunk_len.upload(L.chunk_len.data(), L.chunk_len.size(), s));
        BA_TRY(slot_of.upload(L.slot_of.data(), L.slot_of.size(), s));
        BA_TRY(co_uv64.upload(L.co_uv.data(), L.co_uv.size(), s));
        if (sizeof(Real) != sizeof(double)) {
            std::vector<Real> uvr(L.co_uv.begin(), L.co_uv.end());
            BA_TRY(co_uvr.upload(uvr.data(), uvr.size(), s));
            BA_HIP(hipStreamSynchronize(s));
        }
        // parameters: SoA, points in internal order
        std::vector<double> cx(9 * (size_t)nc), px_(4 * (size_t)np, 0.0);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < 9; ++i) cx[(size_t)i * nc + c] = v->cameras[9 * (size_t)c + i];
        for (int ip = 0; ip < np; ++ip)
            for (int i = 0; i < 3; ++i) px_[4 * (size_t)ip + i] = v->points[3 * (size_t)L.pt_perm[ip] + i];
        BA_TRY(cam_x0.upload(cx.data(), cx.size(), s));
        BA_TRY(pt_x0.upload(px_.data(), px_.size(), s));
        BA_HIP(hipStreamSynchronize(s));
        BA_TRY(cam_x.alloc(9 * (size_t)nc));
        BA_TRY(pt_x.alloc(4 * (size_t)np));
        BA_TRY(cam_c.alloc(9 * (size_t)nc));
        BA_TRY(pt_c.alloc(4 * (size_t)np));
        BA_TRY(dcam.alloc(9 * (size_t)nc));
        BA_TRY(dpt.alloc(4 * (size_t)np));
        BA_TRY(prep.alloc(nc));
        BA_TRY(prep_c.alloc(nc));
        for (DevBuf<double>* b : {&M, &Minv}) BA_TRY(b->alloc(81 * (size_t)nc));
        for (DevBuf<double>* b : {&U, &gc, &sc, &dc2, &bc, &px, &pr, &pz, &pp, &qc}) BA_TRY(b->alloc(9 * (size_t)nc));
        const int n_upd = std::max(1, div_up(std::max(nc, 1), kCamsPerUpd));
        BA_TRY(upd_part.alloc(2 * (size_t)n_upd));
        BA_TRY(rz_hist.alloc((size_t)std::max(o.pcg_max_iterations, 0) + 2));
        BA_TRY(V.alloc(6 * (size_t)np));
        BA_TRY(gp.alloc(3 * (size_t)np));
        BA_TRY(sp.alloc(3 * (size_t)np));
        BA_TRY(dp2.alloc(3 * (size_t)np));
        BA_TRY(Vinv64.alloc(6 * (size_t)np));
        BA_TRY(VinvR.alloc(6 * (size_t)np));
        BA_TRY(qp.alloc(np));
        BA_TRY(zs.alloc(np));
        BA_TRY(co_rec.alloc((size_t)std::max(no, 1)));
        // slot-order buffers: padding slots are never written and must read as zero
        const size_t nslots = (size_t)std::max(L.chunk_off[L.n_chunks], 1);
        BA_TRY(pv.alloc(3 * nslots));
        BA_TRY(u.alloc(nslots));
        BA_HIP(hipMemsetAsync(pv.p, 0, pv.bytes(), s));
        BA_HIP(hipMemsetAsync(u.p, 0, u.bytes(), s));
        const int n_pt_blocks = std::max(1, L.n_chunks);  // point kernels: one block per chunk
        BA_TRY(cam_part.alloc(PS_COUNT * (size_t)std::max(nc, 1)));
        BA_TRY(pt_part.alloc(PS_COUNT * (size_t)n_pt_blocks));
        BA_HIP(hipMemsetAsync(cam_part.p, 0, cam_part.bytes(), s));
        BA_HIP(hipMemsetAsync(pt_part.p, 0, pt_part.bytes(), s));
        BA_TRY(st.alloc(1));
        // one row per trial: a g2o iteration may take g2o_max_trials of them
        const int trace_cap = o.max_num_iterations * (o.mode == BA_MODE_G2O ? std::max(o.g2o_max_trials, 1) : 1) + 2;
        BA_TRY(trace.alloc(trace_cap));
        d.nc = nc;
        d.np = np;
        d.no = no;
        d.n_chunks = L.n_chunks;
        d.n_pt_blocks = n_pt_blocks;
        d.o = make_opts(o);
        d.cam_off = cam_off.p;
        d.co_pt = co_pt.p;
        d.co_uv64 = co_uv64.p;
        d.model = o.mode == BA_MODE_G2O ? kModelSnavelyG2o : kModelSnavely;
        d.co_side = nullptr;
        if (stereo) {
            std::vector<int> side(std::max(no, 1), 0);
            for (int j = 0; j < no; ++j) side[j] = stereo->side[L.co_orig[j]] ? 1 : 0;
            BA_TRY(co_side.upload(side.data(), side.size(), s));
            BA_HIP(hipStreamSynchronize(s));
            d.model = kModelStereo;
            d.co_side = co_side.p;
            d.rig = *stereo->rig;
        }
        d.co_uv = sizeof(Real) == sizeof(double) ? (const Real*)co_uv64.p : co_uvr.p;
        d.co_rec = co_rec.p;

        d.chunk_off = chunk_off.p;
        d.chunk_len = chunk_len.p;
        d.slot_of = slot_of.p;
        d.cam_x = cam_x.p;
        d.pt_x = pt_x.p;
        d.cam_c = cam_c.p;
        d.pt_c = pt_c.p;
        d.dcam = dcam.p;
        d.dpt = dpt.p;
        d.prep = prep.p;
        d.prep_c = prep_c.p;
        d.U = U.p;
        d.gc = gc.p;
        d.sc = sc.p;
        d.dc2 = dc2.p;
        d.qc = qc.p;
        d.dp2 = dp2.p;
        d.upd_part = upd_part.p;
        d.rz_hist = rz_hist.p;
        d.n_upd = n_upd;
        d.M = M.p;
        d.Minv = Minv.p;
        d.bc = bc.p;
        d.px = px.p;
        d.pr = pr.p;
        d.pz = pz.p;
        d.pp = pp.p;
        d.V = V.p;
        d.gp = gp.p;
        d.sp = sp.p;
        d.Vinv64 = Vinv64.p;
        d.VinvR = VinvR.p;
        d.qp = qp.p;
        d.zs = zs.p;
        d.pv = pv.p;
        d.u = u.p;
        d.cam_part = cam_part.p;
        d.pt_part = pt_part.p;
        d.st = st.p;
        d.trace = trace.p;
        d.trace_cap = trace_cap;
        if (linear_solver == BA_SOLVER_DENSE_SCHUR) BA_TRY(dense.create(v, L, o, s));
        BA_TRY(reset());
        return BA_OK;
    }

    int reset() override {
        BA_HIP(hipMemcpyAsync(cam_x.p, cam_x0.p, cam_x0.bytes(), hipMemcpyDeviceToDevice, s));
        BA_HIP(hipMemcpyAsync(pt_x.p, pt_x0.p, pt_x0.bytes(), hipMemcpyDeviceToDevice, s));
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }

    int prep_params(DeviceParams* p) {
        BA_TRY(use_device(device));
        if (!pt_orig.p && L.np) BA_TRY(pt_orig.upload(L.pt_perm.data(), L.pt_perm.size(), s));
        *p = DeviceParams{cam_x.p, pt_x.p, pt_orig.p, L.nc, L.np};
        return BA_OK;
    }
    int normalize() override {
        if (d.model == kModelStereo) return fail(BA_ERR_INVALID_ARGUMENT, "Normalize applies to Snavely cameras only");
        DeviceParams p;
        BA_TRY(prep_params(&p));
        BA_TRY(device_normalize(p, &prep_w, s));
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }
    int perturb(uint32_t seed, double rs, double ts, double ps) override {
        if (d.model == kModelStereo) return fail(BA_ERR_INVALID_ARGUMENT, "Perturb applies to Snavely cameras only");
        if (!(rs >= 0.0) || !(ts >= 0.0) || !(ps >= 0.0)) return fail(BA_ERR_INVALID_ARGUMENT, "negative sigma");
        DeviceParams p;
        BA_TRY(prep_params(&p));
        BA_TRY(device_perturb(p, seed, rs, ts, ps, &prep_w, s));
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }

    int download(double* cams, double* pts) override {
        const int nc = L.nc, np = L.np;
        std::vector<double> cx(9 * (size_t)nc), px_(4 * (size_t)np);
        if (nc) BA_HIP(hipMemcpyAsync(cx.data(), cam_x.p, cx.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        if (np) BA_HIP(hipMemcpyAsync(px_.data(), pt_x.p, px_.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        BA_HIP(hipStreamSynchronize(s));
        if (cams)
            for (int c = 0; c < nc; ++c)
                for (int i = 0; i < 9; ++i) cams[9 * (size_t)c + i] = cx[(size_t)i * nc + c];
        if (pts)
            for (int ip = 0; ip < np; ++ip)
                for (int i = 0; i < 3; ++i) pts[3 * (size_t)L.pt_perm[ip] + i] = px_[4 * (size_t)ip + i];
        return BA_OK;
    }

    // ---- launch helpers ----
    hipEvent_t next_event() {
        if (event_next == event_pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            event_pool.push_back(e);
        }
        return event_pool[event_next++];
    }
    template <typename K, typename... Args>
    void launch_shm(const char* name, K kernel, int grid, int block, size_t shm, Args... args) {
        if (grid <= 0) return;
        hipEvent_t a = nullptr, b = nullptr;
        if (profiling) {
            a = next_event();
            b = next_event();
        }
        if (a && b) {
            // the events are stamped by the dispatch itself: the kernel's own start and end
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (std::uint32_t)shm, s, a, b, 0u, args...);
            recs.push_back({name, a, b});
        } else {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shm, s, args...);
        }
    }
    template <typename K, typename... Args>
    void launch(const char* name, K kernel, int grid, int block, Args... args) {
        launch_shm(name, kernel, grid, block, 0, args...);
    }

    void enqueue_linearize() {
        const int nc = L.nc, gpt = xcd_grid(d.n_pt_blocks), gcam = xcd_grid(nc);
        launch("cam_prep", k_cam_prep<Real>, div_up(nc, 64), 64, d);
        launch("lin_cam", BA_MODEL_KERNEL(k_lin_cam), gcam, kWG, d);
        launch("lin_pt", k_lin_pt<Real>, gpt, kWG, d);
    }

    void enqueue_iteration() {
        const int nc = L.nc, np = L.np;
        const int gpt = xcd_grid(d.n_pt_blocks), gcam = xcd_grid(nc);
        launch("damp_pt", k_damp_pt<Real>, div_up(std::max(np, 1), kWG), kWG, d);
        launch("rhs_cam", BA_MODEL_KERNEL(k_rhs_cam), gcam, kWG, d);
        if (linear_solver == BA_SOLVER_DENSE_SCHUR) {
            dense.enqueue_solve(*this);
        } else {
            launch("cam_factor", k_cam_factor<Real>, d.n_upd, kWG, d);
            // the Snavely model's products in direction form; the g2o and stereo models form J
            const bool dir = d.model == kModelSnavely;
            for (int it = 1; it <= opt.pcg_max_iterations; ++it) {
                if (dir)
                    launch("jtx", k_jtx_dir<Real, kJtxWG>, gcam, kJtxWG, d, it);
                else
                    launch("jtx", BA_MODEL_KERNEL(k_jtx), gcam, kWG, d, 1, it, (const double*)d.pp);
                launch("pt_z", k_pt_z<Real>, gpt, kWG, d);
                if (dir)
                    launch("jz", k_jz_dir<Real, kJzWG>, gcam, kJzWG, d);
                else
                    launch("jz", BA_MODEL_KERNEL(k_jz), gcam, kWG, d);
                launch("cam_update", k_cam_update<Real>, d.n_upd, kWG, d, it);
            }
            // the convergence test after the last update (sets the iteration count)
            if (dir)
                launch("jtx", k_jtx_dir<Real>, gcam, kWG, d, opt.pcg_max_iterations + 1);
            else
                launch("jtx", BA_MODEL_KERNEL(k_jtx), gcam, kWG, d, 1, opt.pcg_max_iterations + 1, (const double*)d.pp);
        }
        launch("jtx_y", BA_MODEL_KERNEL(k_jtx), gcam, kWG, d, 0, 0, (const double*)d.px);
        launch("pt_backsub", k_pt_backsub<Real>, gpt, kWG, d);
        launch("cam_step", BA_MODEL_KERNEL(k_cam_step), div_up(nc, 64), 64, d);
        launch("model_cand", BA_MODEL_KERNEL(k_model_cand), gcam, kWG, d);
        launch("decide", k_decide<Real>, 1, kWG, d);
        launch("apply", k_apply<Real>, div_up(9 * (int64_t)nc + 4 * (int64_t)np, kWG), kWG, d);
        enqueue_linearize();
        launch("post_lin", k_post_lin<Real>, 1, kWG, d);
    }

    // Records one LM iteration (kernel launches only: no memset or copy nodes) into iter_graph.
    // Every argument is fixed for the solver's lifetime (device pointers, PCG iteration numbers);
    // what changes between iterations lives in device memory (the LM state).
    int capture_iteration() {
        hipGraph_t g = nullptr;
        BA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_iteration();
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) return fail(BA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&iter_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            iter_graph = nullptr;
            return fail(BA_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e2));
        }
        return BA_OK;
    }

    int run(ba_summary* summary) override {
        const auto t0 = std::chrono::steady_clock::now();
        BA_TRY(use_device(device));
        recs.clear();
        event_next = 0;
        LMState init{};
        init.radius = opt.initial_trust_region_radius;
        init.decrease_factor = 2.0;
        init.need_lin = 1;
        init.first_lin = 1;
        init.termination = BA_NO_CONVERGENCE;
        *st_host = init;
        BA_HIP(hipMemcpyAsync(st.p, st_host, sizeof(LMState), hipMemcpyHostToDevice, s));
        const int nc = L.nc;
        if (L.no == 0) {
            // nothing to optimise: cost 0, gradient 0
            BA_HIP(hipStreamSynchronize(s));
            if (summary) fill_empty(summary);
            return BA_OK;
        }
        enqueue_linearize();
        launch("prep_x", k_prep_from<Real>, div_up(nc, 64), 64, d, (const double*)d.cam_x);
        launch("cost_x", BA_MODEL_KERNEL(k_cost_cam), xcd_grid(nc), kWG, d, (const double*)d.pt_x);
        launch("post_lin", k_post_lin<Real>, 1, kWG, d);
        BA_HIP(hipGetLastError());
        const int check_every = 4;
        const bool graphs = opt.launch_mode == 0 && !profiling;
        if (graphs && !iter_graph) BA_TRY(capture_iteration());
        // one launch sequence per trial; a g2o iteration takes up to g2o_max_trials of them
        const int max_trials = opt.mode == BA_MODE_G2O ? opt.max_num_iterations * opt.g2o_max_trials : opt.max_num_iterations;
        for (int it = 0; it < max_trials; ++it) {
            if (graphs)
                BA_HIP(hipGraphLaunch(iter_graph, s));
            else
                enqueue_iteration();
            BA_HIP(hipGetLastError());
            if ((it + 1) % check_every == 0 || it + 1 == max_trials) {
                BA_HIP(hipMemcpyAsync(st_host, st.p, sizeof(LMState), hipMemcpyDeviceToHost, s));
                BA_TRY(sync("LM state"));
                if (st_host->done) break;
            }
        }
        BA_HIP(hipMemcpyAsync(st_host, st.p, sizeof(LMState), hipMemcpyDeviceToHost, s));
        BA_TRY(sync("LM state"));
        const auto t1 = std::chrono::steady_clock::now();
        if (profiling) collect_profile();
        if (summary) BA_TRY(fill_summary(summary, std::chrono::duration<double, std::milli>(t1 - t0).count()));
        if (opt.verbose) print_trace();
        return BA_OK;
    }

    void fill_empty(ba_summary* sm) {
        sm->initial_cost = sm->final_cost = 0.0;
        sm->num_iterations = 0;
        sm->num_successful_steps = sm->num_unsuccessful_steps = 0;
        sm->termination_type = BA_CONVERGENCE;
        sm->termination_reason = BA_REASON_NONE;
        sm->termination_value = 0.0;
        sm->linear_solver_used = linear_solver;
        sm->total_linear_solver_iterations = 0;
        sm->total_time_ms = 0.0;
    }

    int fill_summary(ba_summary* sm, double ms) {
        const LMState& S = *st_host;
        const int n = std::min(S.trace_len, d.trace_cap);
        std::vector<ba_iteration> rows(n);
        if (n) {
            BA_HIP(hipMemcpyAsync(rows.data(), trace.p, n * sizeof(ba_iteration), hipMemcpyDeviceToHost, s));
            BA_TRY(sync("LM trace"));
        }
        if (opt.mode == BA_MODE_G2O)
            for (ba_iteration& r : rows) {  // g2o reports chi2 = sum rho(s), without the 1/2
                r.cost *= 2.0;
                r.cost_change *= 2.0;
                r.model_cost_change *= 2.0;
            }
        // device clock ticks since the first evaluation -> seconds; per-row deltas
        double prev = 0.0;
        for (ba_iteration& r : rows) {
            r.cumulative_time_in_seconds /= clock_khz * 1e3;
            r.iteration_time_in_seconds = r.cumulative_time_in_seconds - prev;
            prev = r.cumulative_time_in_seconds;
        }
        const double cscale = opt.mode == BA_MODE_G2O ? 2.0 : 1.0;
        sm->initial_cost = n ? rows[0].cost : cscale * S.x_cost;
        sm->final_cost = cscale * S.x_cost;
        sm->num_iterations = S.trace_len;
        sm->num_successful_steps = S.num_successful;
        sm->num_unsuccessful_steps = S.num_unsuccessful;
        sm->termination_type = S.termination;
        sm->termination_reason = S.reason;
        sm->termination_value = S.reason_value;
        sm->linear_solver_used = linear_solver;
        sm->total_linear_solver_iterations = S.total_pcg_iters;
        sm->total_time_ms = ms;
        if (sm->trace && sm->trace_capacity > 0) {
            const int m = std::min(n, sm->trace_capacity);
            std::copy(rows.begin(), rows.begin() + m, sm->trace);
        }
        last_rows = rows;
        return BA_OK;
    }

    std::vector<ba_iteration> last_rows;

    void print_trace() {
        printf("iter      cost      cost_change  |gradient|   |step|    tr_ratio  tr_radius  ls_iter  iter_time  "
               "total_time\n");
        for (const ba_iteration& r : last_rows)
            printf("%4d % 8.6e % 3.2e % 3.2e % 3.2e % 3.2e % 3.2e %4d % 3.2e % 3.2e\n", r.iteration, r.cost,
                   r.cost_change, r.gradient_max_norm, r.step_norm, r.relative_decrease, r.trust_region_radius,
                   r.linear_solver_iterations, r.iteration_time_in_seconds, r.cumulative_time_in_seconds);
        fflush(stdout);
    }

    // ---- profiling ----
    int set_profiling(bool on) override {
        profiling = on;
        ktimes.clear();
        return BA_OK;
    }

    double bytes_of(const char* name) const {
        const double R = sizeof(Real), no = L.no, np = L.np, nc = L.nc;
        const std::string n(name);
        // Algorithmic bytes of one launch: every stream counted once per pass, per-camera and per-point
        // parameters once per pass, FP64 parameters, int32 indices.
        const bool dir = d.model == kModelSnavely && linear_solver == BA_SOLVER_ITERATIVE_SCHUR;
        // direction form: the observation record (X - c, Huber weight), slot, u out (jtx); the record,
        // point index and the point's FP64 zs (jz); per camera its prep and the 9-vectors read and
        // written
        const double prep_b = sizeof(CameraPrep<Real>);
        if (n == "jtx" && dir) return no * (4 * R + 4 + 4 * R) + nc * (prep_b + 9 * 8 * 4 + 2 * 8 * d.n_upd);
        if (n == "jz" && dir) return no * (4 * R + 4) + np * 32 + nc * (prep_b + 9 * 8 * 5);
        if (n == "jtx" || n == "jtx_y") return no * (4 + 4 + 2 * R + 4 * R) + np * 4 * 8 + nc * 9 * 8 * 3;
        if (n == "jz") return no * (4 + 2 * R) + np * (4 * 8 + 4 * R) + nc * 9 * 8 * 8;
        // u by slot; per point its scale, FP64 V*^-1 (6), D_p^2, and the FP64 zs written
        if (n == "pt_z") return no * 4 * R + np * (3 * 8 + 6 * 8 + 3 * 8 + 32);
        if (n == "lin_cam") return no * (4 + 4 + 2 * R + 12 * 8) + np * 4 * 8 + nc * 18 * 8;
        if (n == "lin_pt") return no * 12 * 8 + np * 9 * 8;
        if (n == "rhs_cam") return no * (4 + 2 * R) + np * (3 * 8 + 3 * 8 + 6 * R + 3 * R) + nc * (81 * 3 + 9 * 3) * 8;
        if (n == "model_cand") return no * (4 + 2 * R + 16) + np * 3 * 8 * 3 + nc * 9 * 8;
        return 0.0;
    }

    void collect_profile() {
        (void)hipStreamSynchronize(s);
        for (const Rec& r : recs) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
            auto it = std::find_if(ktimes.begin(), ktimes.end(), [&](const ba_kernel_time& k) { return strcmp(k.name, r.name) == 0; });
            if (it == ktimes.end()) {
                ba_kernel_time k{};
                snprintf(k.name, sizeof(k.name), "%s", r.name);
                k.bytes_per_launch = bytes_of(r.name);
                ktimes.push_back(k);
                it = ktimes.end() - 1;
            }
            it->launches++;
            it->total_ms += ms;
        }
        recs.clear();
        event_next = 0;
    }

    int kernel_times(ba_kernel_time* out, int cap, int* count) override {
        const int n = (int)ktimes.size();
        if (count) *count = n;
        for (int i = 0; i < std::min(n, cap); ++i) out[i] = ktimes[i];
        return BA_OK;
    }
};

int make_solver(const ba_problem_view* v, const ba_options& o, std::unique_ptr<SolverBase>* out,
                const StereoInput* stereo) {
    if (o.precision == BA_PRECISION_FP64) {
        auto s = std::make_unique<Solver<double>>();
        BA_TRY(s->create(v, o, stereo));
        *out = std::move(s);
    } else {
        auto s = std::make_unique<Solver<float>>();
        BA_TRY(s->create(v, o, stereo));
        *out = std::move(s);
    }
    return BA_OK;
}

}  // namespace mba

// =============================================================================================
struct ba_solver {
    std::unique_ptr<mba::SolverBase> impl;
    ba_problem_view view{};
};

using namespace mba;

extern "C" {

int ba_solve(const ba_problem_view* v, const ba_options* options, ba_summary* summary) {
    clear_error();
    ba_options o;
    if (options)
        o = *options;
    else
        ba_default_options(&o);
    BA_TRY(validate_view(v, true));
    BA_TRY(check_options(&o));
    std::unique_ptr<SolverBase> s;
    BA_TRY(make_solver(v, o, &s));
    BA_TRY(s->run(summary));
    return s->download(v->cameras, v->points);
}

void ba_stereo_default_options(ba_options* o) {
    if (!o) return;
    ba_default_options(o);
    o->mode = BA_MODE_G2O;                       // OptimizationAlgorithmLevenberg (backend.cpp:71-75)
    o->max_num_iterations = 10;                  // optimizer.optimize(10) (backend.cpp:150)
    o->loss_scale = 5.991;                       // RobustKernelHuber delta = chi2_th (backend.cpp:102,139-141)
    o->linear_solver = BA_SOLVER_DENSE_SCHUR;    // LinearSolverCSparse on the 6-dof pose blocks
}

}  // extern "C"

namespace mba {

int stereo_prepare(const ba_stereo_view* v, StereoProblem* out) {
    if (!v) return fail(BA_ERR_INVALID_ARGUMENT, "problem view is NULL");
    if (v->num_observations > 0 && !v->side) return fail(BA_ERR_INVALID_ARGUMENT, "side array is NULL");
    for (int k = 0; k < v->num_observations; ++k)
        if (v->side[k] != 0 && v->side[k] != 1)
            return fail(BA_ERR_INVALID_ARGUMENT, "observation %d: side must be 0 (left) or 1 (right)", k);
    for (int i = 0; i < 4; ++i)
        if (!std::isfinite(v->K[i])) return fail(BA_ERR_NON_FINITE, "K is not finite");
    if (v->K[0] == 0.0 || v->K[1] == 0.0) return fail(BA_ERR_INVALID_ARGUMENT, "focal lengths must be nonzero");
    if (v->num_poses > 0 && !v->poses) return fail(BA_ERR_INVALID_ARGUMENT, "poses is NULL");
    // poses in the solver's 9-parameter slots (angle-axis, t, three unused entries)
    out->cams.assign(9 * (size_t)std::max(v->num_poses, 0), 0.0);
    for (int i = 0; i < v->num_poses; ++i)
        for (int k = 0; k < 6; ++k) out->cams[9 * (size_t)i + k] = v->poses[6 * (size_t)i + k];
    out->view = ba_problem_view{v->num_poses, v->num_points, v->num_observations, 9, v->pose_index, v->point_index,
                                v->observations, out->cams.data(), v->points};
    BA_TRY(validate_view(&out->view, true));
    out->rig = StereoRig{};
    for (int i = 0; i < 4; ++i) out->rig.K[i] = v->K[i];
    for (int sd = 0; sd < 2; ++sd) {
        double e9[9] = {0};
        for (int k = 0; k < 6; ++k) e9[k] = v->extrinsics[sd][k];
        CameraPrep<double> P;
        camera_prep<double>(e9, P);
        for (int k = 0; k < 9; ++k) out->rig.R[sd][k] = P.R64[k];
        for (int k = 0; k < 3; ++k) out->rig.t[sd][k] = P.t64[k];
    }
    out->input = StereoInput{&out->rig, v->side};
    return BA_OK;
}

void stereo_finish(const ba_stereo_view* v, const StereoProblem& sp, double chi2_start, uint8_t* outlier,
                   double* chi2_threshold) {
    for (int i = 0; i < v->num_poses; ++i)
        for (int k = 0; k < 6; ++k) v->poses[6 * (size_t)i + k] = sp.cams[9 * (size_t)i + k];
    // the outlier rule of backend.cpp:153-186 on the optimised parameters
    const int no = v->num_observations;
    std::vector<double> chi2(no);
    for (int k = 0; k < no; ++k) {
        CameraPrep<double> P;
        camera_prep<double>(&sp.cams[9 * (size_t)v->pose_index[k]], P);
        double r[2];
        stereo_residual<double>(P, &v->points[3 * (size_t)v->point_index[k]], sp.rig, v->side[k],
                                v->observations[2 * (size_t)k], v->observations[2 * (size_t)k + 1], r);
        chi2[k] = r[0] * r[0] + r[1] * r[1];
    }
    double th = chi2_start;
    for (int it = 0; it < 5; ++it) {
        int in_n = 0;
        for (int k = 0; k < no; ++k) in_n += chi2[k] > th ? 0 : 1;
        if (in_n / (double)no > 0.5) break;  // (no observations: 0/0, never > 0.5, as in the reference)
        th *= 2.0;
    }
    if (outlier)
        for (int k = 0; k < no; ++k) outlier[k] = chi2[k] > th ? 1 : 0;
    if (chi2_threshold) *chi2_threshold = th;
}

}  // namespace mba

extern "C" {

int ba_stereo_solve(const ba_stereo_view* v, const ba_options* options, ba_summary* summary, uint8_t* outlier,
                    double* chi2_threshold) {
    clear_error();
    ba_options o;
    if (options)
        o = *options;
    else
        ba_stereo_default_options(&o);
    BA_TRY(check_options(&o));
    StereoProblem sp;
    BA_TRY(stereo_prepare(v, &sp));
    std::unique_ptr<SolverBase> sv;
    BA_TRY(make_solver(&sp.view, o, &sv, &sp.input));
    BA_TRY(sv->run(summary));
    BA_TRY(sv->download(sp.cams.data(), v->points));
    stereo_finish(v, sp, o.loss_scale, outlier, chi2_threshold);
    return BA_OK;
}

int ba_solver_create(const ba_problem_view* v, const ba_options* options, ba_solver** out) {
    clear_error();
    if (!out) return fail(BA_ERR_INVALID_ARGUMENT, "out is NULL");
    ba_options o;
    if (options)
        o = *options;
    else
        ba_default_options(&o);
    BA_TRY(validate_view(v, true));
    BA_TRY(check_options(&o));
    auto h = std::make_unique<ba_solver>();
    BA_TRY(make_solver(v, o, &h->impl));
    h->view = *v;
    *out = h.release();
    return BA_OK;
}

int ba_solver_run(ba_solver* h, ba_summary* summary) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->run(summary);
}

int ba_solver_reset(ba_solver* h) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->reset();
}

int ba_solver_download(ba_solver* h, double* cameras, double* points) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->download(cameras, points);
}

int ba_solver_normalize(ba_solver* h) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->normalize();
}

int ba_solver_perturb(ba_solver* h, uint32_t seed, double rotation_sigma, double translation_sigma,
                      double point_sigma) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->perturb(seed, rotation_sigma, translation_sigma, point_sigma);
}

int ba_solver_destroy(ba_solver* h) {
    delete h;
    return BA_OK;
}

int ba_solver_set_profiling(ba_solver* h, int32_t enabled) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    return h->impl->set_profiling(enabled != 0);
}

int ba_solver_kernel_times(ba_solver* h, ba_kernel_time* out, int32_t capacity, int32_t* count) {
    clear_error();
    if (!h || (!out && capacity > 0)) return fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    int n = 0;
    BA_TRY(h->impl->kernel_times(out, capacity, &n));
    if (count) *count = n;
    return BA_OK;
}

int ba_solver_info(ba_solver* h, int64_t* device_bytes, int32_t* linear_solver) {
    clear_error();
    if (!h) return fail(BA_ERR_INVALID_ARGUMENT, "NULL solver");
    if (device_bytes) *device_bytes = h->impl->device_bytes();
    if (linear_solver) *linear_solver = h->impl->solver_kind();
    return BA_OK;
}

}  // extern "C"
