// window.hip -- batched local bundle adjustment (ba_batch_*) and the serving loop (ba_server_*).
//
// Windows (small BA problems: the local BA that full_slam/src/backend.cpp:67-199 runs over its
// 7-10 active keyframes, here with the Snavely camera of the BA path) are packed into regions of
// one device arena; one LM iteration of every active window is the fixed sequence of phase
// kernels of window_kernels.h (captured once and replayed as a hipGraph for a fixed batch).
// Windows whose reduced camera system does not fit a workgroup's LDS are solved by the
// single-problem solver.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "solver.h"
#include "window_kernels.h"

namespace mba {

namespace {

// LDS a workgroup may allocate on the current device (160 KiB on MI355X), queried once.
int64_t lds_per_block() {
    static int64_t cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            cached = v;
        else
            cached = 65536;
    }
    return cached;
}

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }


// Static LDS of k_wsolve (xv, colv, dinv, bad) plus slack for alignment.
int64_t solve_static_lds() { return 3 * kWinMaxN * 8 + 16 + 256; }

// Largest reduced system (9 K) k_wsolve can factor in LDS (0: none).
int max_window_n() {
    const int64_t room = lds_per_block() - solve_static_lds();
    if (room <= 0) return 0;
    int n = (int)std::floor(std::sqrt(2.0 * (double)room / 8.0));  // packed: n (n + 1) / 2 doubles
    while (n > 0 && (int64_t)n * (n + 1) / 2 * 8 > room) --n;
    n = std::min(n, kWinMaxN);
    return n - n % 9;
}

// Dynamic LDS of k_wschur: one staging area per wave.
template <typename Real>
int64_t schur_lds(int nt) {
    return (int64_t)(nt / 64) * win_stage_bytes<Real>();
}

}  // namespace

// Host image of one window's arena region (offsets relative to the region start).
struct WindowImage {
    WinDesc desc{};
    std::vector<char> bytes;  // the region's inputs (the first bytes.size() bytes of the region)
    int64_t total = 0;        // bytes of the whole region; the rest starts zeroed on the device
    std::vector<int> pt_perm;
    int K = 0, P = 0;
    int64_t cost = 0;  // work estimate for dispatch order
};

// stereo: the full_slam model (rig and per-observation camera side, caller order); NULL for the
// Snavely model.
template <typename Real>
int build_window_image(const ba_problem_view* v, const ba_options& opt, WindowImage* img,
                       const StereoInput* stereo = nullptr) {
    Layout L;
    BA_TRY(build_layout(v, &L));
    PairBlocks pb;
    build_pair_blocks(L, &pb);
    const int K = L.nc, P = L.np, No = L.no, nch = L.n_chunks;
    const int slots = L.chunk_off[nch];
    // a diagonal block whose pairs are all an observation with itself is formed by k_wrec_cam; the
    // pair kernel takes the blocks of two different cameras, and the diagonal blocks of a camera
    // that observes one point more than once (their cross pairs x != y carry the terms
    // -J~c,x' J~p,x V*^-1 J~p,y' J~c,y that k_wrec_cam does not form; it then rewrites the whole
    // block after k_wrec_cam)
    std::vector<int> offdiag;
    for (int i = 0; i < (int)pb.a.size(); ++i) {
        bool take = pb.a[i] != pb.b[i];
        for (int q = pb.off[i]; !take && q < pb.off[i + 1]; ++q) take = pb.i[q] != pb.j[q];
        if (take) offdiag.push_back(i);
    }
    const int nb = (int)offdiag.size(), npairs = (int)pb.i.size();
    const int n = 9 * K;
    WinDesc& D = img->desc;
    D = WinDesc{};
    D.K = K;
    D.P = P;
    D.No = No;
    D.n_chunks = nch;
    D.n_slots = slots;
    D.n_blocks = nb;
    D.n_pairs = npairs;
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        const int64_t o = off;
        off = align256(off + std::max<int64_t>(bytes, 1));
        return o;
    };
    D.o_cam_off = take(4 * (K + 1));
    D.o_co_pt = take(4 * (int64_t)No);
    D.o_co_uv64 = take(16 * (int64_t)No);
    D.o_slot_of = take(4 * (int64_t)No);
    D.o_chunk_len = take(4 * (int64_t)nch);
    D.o_chunk_off = take(4 * ((int64_t)nch + 1));
    D.o_blk = take(16 * (int64_t)nb);
    D.o_pairs = take(8 * (int64_t)npairs);
    D.o_cam_x = take(72 * (int64_t)K);
    D.o_cam_x0 = take(72 * (int64_t)K);
    D.o_pt_x = take(32 * (int64_t)P);
    D.o_pt_x0 = take(32 * (int64_t)P);
    D.o_side = take(stereo ? 4 * (int64_t)No : 0);
    D.o_rig = take(stereo ? (int64_t)sizeof(StereoRig) : 0);
    D.o_state = take(sizeof(WinState));
    const int64_t input_bytes = off;  // everything below is work space, zeroed on the device
    D.o_cam_c = take(72 * (int64_t)K);
    D.o_pt_c = take(32 * (int64_t)P);
    D.o_prep = take(sizeof(CameraPrep<double>) * (int64_t)K);
    D.o_prep_c = take(sizeof(CameraPrep<double>) * (int64_t)K);
    D.o_udiag = take(72 * (int64_t)K);
    D.o_gc = take(72 * (int64_t)K);
    D.o_sc = take(72 * (int64_t)K);
    D.o_dc2 = take(72 * (int64_t)K);
    D.o_bc = take(72 * (int64_t)K);
    D.o_yc = take(72 * (int64_t)K);
    D.o_cpart = take(8 * WC_COUNT * (int64_t)K);
    D.o_ppart = take(8 * WP_COUNT * (int64_t)nch);
    D.o_V = take(48 * (int64_t)P);
    D.o_gp = take(24 * (int64_t)P);
    D.o_sp = take(24 * (int64_t)P);
    D.o_Vinv = take(48 * (int64_t)P);
    D.o_qp = take(32 * (int64_t)P);
    D.o_yp = take(24 * (int64_t)P);
    D.o_pv = take(96 * (int64_t)slots);
    D.o_js = take(kWinJs * sizeof(Real) * (int64_t)No);
    D.o_S = take(sizeof(double) * (int64_t)n * n);  // zero: blocks without common points are never written
    D.o_L = take(sizeof(double) * (int64_t)n * (n + 1) / 2);  // the factor when it is not kept in LDS
    img->total = off;
    img->bytes.assign(input_bytes, 0);
    char* b = img->bytes.data();
    std::memcpy(b + D.o_cam_off, L.cam_off.data(), 4 * (size_t)(K + 1));
    std::memcpy(b + D.o_co_pt, L.co_pt.data(), 4 * (size_t)No);
    std::memcpy(b + D.o_co_uv64, L.co_uv.data(), 16 * (size_t)No);
    std::memcpy(b + D.o_slot_of, L.slot_of.data(), 4 * (size_t)No);
    std::memcpy(b + D.o_chunk_len, L.chunk_len.data(), 4 * (size_t)nch);
    std::memcpy(b + D.o_chunk_off, L.chunk_off.data(), 4 * (size_t)(nch + 1));
    // blocks with the most pairs first: the workgroup's waves take them in this order
    std::vector<int> border(offdiag);
    std::stable_sort(border.begin(), border.end(),
                     [&](int x, int y) { return pb.off[x + 1] - pb.off[x] > pb.off[y + 1] - pb.off[y]; });
    int* blk = reinterpret_cast<int*>(b + D.o_blk);
    for (int k = 0; k < nb; ++k) {
        const int i = border[k];
        blk[4 * k] = pb.a[i];
        blk[4 * k + 1] = pb.b[i];
        blk[4 * k + 2] = pb.off[i];
        blk[4 * k + 3] = pb.off[i + 1];
    }
    int* pr = reinterpret_cast<int*>(b + D.o_pairs);
    for (int i = 0; i < npairs; ++i) {
        pr[2 * i] = pb.i[i];
        pr[2 * i + 1] = pb.j[i];
    }
    std::memcpy(b + D.o_cam_x, v->cameras, 72 * (size_t)K);
    std::memcpy(b + D.o_cam_x0, v->cameras, 72 * (size_t)K);
    if (stereo) {
        int* sd = reinterpret_cast<int*>(b + D.o_side);
        for (int j = 0; j < No; ++j) sd[j] = stereo->side[L.co_orig[j]] ? 1 : 0;
        std::memcpy(b + D.o_rig, stereo->rig, sizeof(StereoRig));
    }
    double* px = reinterpret_cast<double*>(b + D.o_pt_x);
    double* px0 = reinterpret_cast<double*>(b + D.o_pt_x0);
    for (int ip = 0; ip < P; ++ip)
        for (int i = 0; i < 3; ++i) px[4 * (size_t)ip + i] = px0[4 * (size_t)ip + i] = v->points[3 * (size_t)L.pt_perm[ip] + i];
    WinState* st = reinterpret_cast<WinState*>(b + D.o_state);
    *st = WinState{};
    st->radius = opt.initial_trust_region_radius;
    st->decrease_factor = 2.0;
    st->need_lin = 1;
    st->first_lin = 1;
    st->termination = BA_NO_CONVERGENCE;
    if (No == 0) {
        st->done = 1;
        st->termination = BA_CONVERGENCE;
        st->first_lin = 0;
    }
    img->pt_perm = L.pt_perm;
    img->K = K;
    img->P = P;
    img->cost = (int64_t)No * 16 + (int64_t)npairs * 4 + (int64_t)n * n * n / 32;
    return BA_OK;
}

WinDesc rebase(const WinDesc& d, int64_t base) {
    WinDesc r = d;
    int64_t* o = &r.o_cam_off;
    const int count = (int)((&r.o_state - &r.o_cam_off) + 1);
    for (int i = 0; i < count; ++i) o[i] += base;
    return r;
}

// Restores a window's initial parameters and LM state.
__global__ void k_window_reset(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                               const int* __restrict__ work, int n_work, WinState init) {
    const int wi = blockIdx.x;
    if (wi >= n_work) return;
    const WinDesc D = descs[work[wi]];
    double* cx = at<double>(arena, D.o_cam_x);
    const double* cx0 = at<double>(arena, D.o_cam_x0);
    double4* px = at<double4>(arena, D.o_pt_x);
    const double4* px0 = at<double4>(arena, D.o_pt_x0);
    for (int k = threadIdx.x; k < 9 * D.K; k += blockDim.x) cx[k] = cx0[k];
    for (int p = threadIdx.x; p < D.P; p += blockDim.x) px[p] = px0[p];
    if (threadIdx.x == 0) {
        WinState s = init;
        if (D.No == 0) {
            s.done = 1;
            s.termination = BA_CONVERGENCE;
            s.first_lin = 0;
        }
        *at<WinState>(arena, D.o_state) = s;
    }
}

// Gathers the LM states of the listed windows into out (one thread per window).
__global__ void k_window_states(const WinDesc* __restrict__ descs, const char* __restrict__ arena,
                                const int* __restrict__ work, int n_work, WinState* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_work) return;
    out[i] = *reinterpret_cast<const WinState*>(arena + descs[work[i]].o_state);
}

// ---------------------------------------------------------------------------------------------
// The engine: an arena of window regions, slot descriptors, launches.
template <typename Real>
struct WindowEngine {
    ba_options opt{};
    Opts o{};
    hipStream_t s = nullptr;
    DevBuf<char> arena;
    int64_t arena_cap = 0;
    DevBuf<WinDesc> ddesc;
    std::vector<WinDesc> hdesc;  // absolute offsets, by slot
    DevBuf<int> dwork;
    DevBuf<WinState> dstates;
    std::vector<WinState> hstates;
    // schedule (see WinSchedule): none of these changes a window's result
    int threads = 256;        // block size of k_wschur
    int solve_threads = 256;  // block size of k_wsolve
    int blocks_per_wave = 2;  // 9x9 blocks of S per wave of k_wschur (task granularity)
    int lds_cameras = kWinMaxK;  // windows of at most this many cameras factor in LDS
    int model = kModelSnavely;   // residual model of every window (kModelSnavely or kModelStereo)
    int max_slots = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the current work set: task lists of the phase kernels (windows in dispatch order)
    std::vector<int> work_slots;
    DevBuf<int2> ctask, ptask;
    DevBuf<int4> stask;
    DevBuf<int> wtask;
    int n_ctask = 0, n_ptask = 0, n_stask = 0, n_wtask = 0, work_nmax = 9, work_threads = 0, work_solve = 0,
        work_bpw = 0, work_lds = 0;
    hipGraphExec_t graph = nullptr;  // one LM iteration of the current work set
    double wait_timeout_s = 0.0;     // > 0: device waits give up with BA_ERR_TIMEOUT after this long
    hipEvent_t ev_wait = nullptr;
    DevBuf<unsigned long long> dclock;  // device wall clock at the start of the last run
    double clock_khz = 100000.0;        // rate of the device wall clock

    bool own_stream = true;  // false: the stream is a lane stream of the pool (not destroyed here)

    ~WindowEngine() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_wait) (void)hipEventDestroy(ev_wait);
        if (s && own_stream) (void)hipStreamDestroy(s);
    }

    // Waits for the engine stream (bounded when wait_timeout_s > 0).
    int sync(const char* what) { return wait_stream(s, ev_wait, wait_timeout_s, what); }

    // stream: a lane stream to launch on (kept by its owner), or NULL for a stream of the engine's own
    int init(const ba_options& options, int64_t arena_bytes, int slots, hipStream_t stream = nullptr) {
        opt = options;
        o = solver_opts(options);
        BA_TRY(use_device(options.device));
        if (stream) {
            s = stream;
            own_stream = false;
        } else {
            BA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        }
        BA_HIP(hipEventCreate(&ev0));
        BA_HIP(hipEventCreate(&ev1));
        BA_HIP(hipEventCreateWithFlags(&ev_wait, hipEventDisableTiming));
        arena_cap = std::max<int64_t>(arena_bytes, 256);
        BA_TRY(arena.alloc((size_t)arena_cap));
        max_slots = std::max(slots, 1);
        BA_TRY(ddesc.alloc(max_slots));
        hdesc.assign(max_slots, WinDesc{});
        BA_TRY(dwork.alloc(max_slots));
        BA_TRY(dstates.alloc(max_slots));
        BA_TRY(dclock.alloc(1));
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            clock_khz = khz;
        return BA_OK;
    }

    // Device wall clock at the start of the last run.
    int clock_start(unsigned long long* t0) {
        BA_HIP(hipMemcpyAsync(t0, dclock.p, sizeof(*t0), hipMemcpyDeviceToHost, s));
        return sync("window clock");
    }

    // Milliseconds from the start of the last run to the window's finish (0 if it has not finished).
    int finish_times(const std::vector<WinState>& st, std::vector<double>* ms) {
        unsigned long long t0 = 0;
        BA_HIP(hipMemcpyAsync(&t0, dclock.p, sizeof(t0), hipMemcpyDeviceToHost, s));
        BA_TRY(sync("window clock"));
        ms->assign(st.size(), 0.0);
        for (size_t i = 0; i < st.size(); ++i)
            if (st[i].done && st[i].t_done >= t0) (*ms)[i] = (double)(st[i].t_done - t0) / clock_khz;
        return BA_OK;
    }

    WinState init_state() const {
        WinState st{};
        st.radius = opt.initial_trust_region_radius;
        st.decrease_factor = 2.0;
        st.need_lin = 1;
        st.first_lin = 1;
        st.termination = BA_NO_CONVERGENCE;
        return st;
    }

    // Places an image at arena offset `base` in slot `slot` (asynchronous on the engine stream;
    // the image must stay alive until the stream has consumed it).
    int place(const WindowImage& img, int slot, int64_t base) {
        if (base + img.total > arena_cap) return fail(BA_ERR_CAPACITY, "window does not fit the arena");
        const size_t in = img.bytes.size();
        BA_HIP(hipMemcpyAsync(arena.p + base, img.bytes.data(), in, hipMemcpyHostToDevice, s));
        BA_HIP(hipMemsetAsync(arena.p + base + in, 0, (size_t)img.total - in, s));
        hdesc[slot] = rebase(img.desc, base);
        BA_HIP(hipMemcpyAsync(ddesc.p + slot, &hdesc[slot], sizeof(WinDesc), hipMemcpyHostToDevice, s));
        work_slots.clear();  // the task lists depend on the windows' sizes: rebuild them
        return BA_OK;
    }

    // Reduced systems of at most lds_n() unknowns are factored in LDS (the launch reserves the
    // largest such window's triangle for every workgroup), larger ones in their global buffer.
    int lds_n() const { return std::min(9 * lds_cameras, max_window_n()); }

    template <typename T>
    int upload(DevBuf<T>& buf, const std::vector<T>& v) {
        if (v.empty()) return BA_OK;
        if (buf.n < v.size()) BA_TRY(buf.alloc(v.size()));
        BA_HIP(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        return BA_OK;
    }

    // Task lists of the phase kernels for the windows in `slots` (in this dispatch order).
    int set_work(const std::vector<int>& slots) {
        if (slots == work_slots && work_threads == threads && work_solve == solve_threads &&
            work_bpw == blocks_per_wave && work_lds == lds_n())
            return BA_OK;
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        std::vector<int2> ct, pt;
        const int nw = threads / 64, group = blocks_per_wave * nw;  // blocks per k_wschur task
        work_nmax = 9;
        // k_wschur: every block group of a window goes to the same XCD (workgroup k runs on XCD
        // k mod 8), so the window's records, read once per pair, are re-read from that XCD's L2;
        // windows are spread over the XCDs largest first onto the least loaded one
        std::vector<std::vector<int4>> xq(kXcds);
        std::vector<double> xload(kXcds, 0.0);
        for (int sl : slots) {
            const WinDesc& D = hdesc[sl];
            if (9 * D.K <= lds_n()) work_nmax = std::max(work_nmax, 9 * D.K);
            for (int c = 0; c < D.K; ++c) ct.push_back(make_int2(sl, c));
            for (int ch = 0; ch < D.n_chunks; ++ch) pt.push_back(make_int2(sl, ch));
            const int x = (int)(std::min_element(xload.begin(), xload.end()) - xload.begin());
            xload[x] += (double)D.n_pairs + 1.0;
            const int G = std::max(1, (D.n_blocks + group - 1) / group);
            for (int g = 0; g < G && g < D.n_blocks; ++g)
                xq[x].push_back(make_int4(sl, g, G, (D.n_blocks - g + G - 1) / G));
        }
        size_t qmax = 0;
        for (const auto& q : xq) qmax = std::max(qmax, q.size());
        std::vector<int4> stk(qmax * kXcds);
        for (size_t k = 0; k < qmax; ++k)
            for (int x = 0; x < kXcds; ++x)  // a padding task names a block group past the window's last
                stk[k * kXcds + x] = k < xq[x].size() ? xq[x][k] : make_int4(slots[0], 1 << 30, 1, 1);
        BA_TRY(upload(ctask, ct));
        BA_TRY(upload(ptask, pt));
        BA_TRY(upload(stask, stk));
        BA_TRY(upload(wtask, slots));
        BA_TRY(sync("window task lists"));
        n_ctask = (int)ct.size();
        n_ptask = (int)pt.size();
        n_stask = (int)stk.size();
        n_wtask = (int)slots.size();
        work_slots = slots;
        work_threads = threads;
        work_solve = solve_threads;
        work_bpw = blocks_per_wave;
        work_lds = lds_n();
        return BA_OK;
    }

    // M: residual model, NT: block size of k_wschur, NF: of k_wsolve
    template <int M, int NT, int NF>
    int enqueue_iteration_nt() {
        const WinDesc* dd = ddesc.p;
        if (n_ctask) hipLaunchKernelGGL((k_wlin_cam<kWinCamThreads, M>), dim3(n_ctask), dim3(kWinCamThreads), 0, s, dd,
                                        arena.p, (const int2*)ctask.p, n_ctask, o);
        if (n_ptask) hipLaunchKernelGGL(k_wlin_pt, dim3(n_ptask), dim3(64), 0, s, dd, arena.p, (const int2*)ptask.p,
                                        n_ptask, o);
        hipLaunchKernelGGL((k_wdamp<256>), dim3(n_wtask), dim3(256), 0, s, dd, arena.p, (const int*)wtask.p, n_wtask, o);
        if (n_ctask) hipLaunchKernelGGL((k_wrec_cam<Real, kWinCamThreads, M>), dim3(n_ctask), dim3(kWinCamThreads), 0, s,
                                        dd, arena.p, (const int2*)ctask.p, n_ctask, o);
        if (n_stask)
            hipLaunchKernelGGL((k_wschur<Real, NT>), dim3(n_stask), dim3(NT), (size_t)schur_lds<Real>(NT), s, dd, arena.p,
                               (const int4*)stask.p, n_stask);
        hipLaunchKernelGGL((k_wsolve<NF>), dim3(n_wtask), dim3(NF), solve_lds(), s, dd, arena.p, (const int*)wtask.p,
                           n_wtask, lds_n());
        if (n_ctask) hipLaunchKernelGGL((k_wback_cam<kWinCamThreads, M>), dim3(n_ctask), dim3(kWinCamThreads), 0, s, dd,
                                        arena.p, (const int2*)ctask.p, n_ctask, o);
        if (n_ptask) hipLaunchKernelGGL(k_wback_pt, dim3(n_ptask), dim3(64), 0, s, dd, arena.p, (const int2*)ptask.p,
                                        n_ptask, o);
        if (n_ctask) hipLaunchKernelGGL((k_wcand_cam<kWinCamThreads, M>), dim3(n_ctask), dim3(kWinCamThreads), 0, s, dd,
                                        arena.p, (const int2*)ctask.p, n_ctask, o);
        hipLaunchKernelGGL((k_wdecide<256>), dim3(n_wtask), dim3(256), 0, s, dd, arena.p, (const int*)wtask.p, n_wtask,
                           o);
        return BA_OK;
    }

    int set_lds_limits() {
        if (threads == 512)
            BA_TRY(allow_dynamic_lds((const void*)k_wschur<Real, 512>, (int)schur_lds<Real>(512)));
        else
            BA_TRY(allow_dynamic_lds((const void*)k_wschur<Real, 256>, (int)schur_lds<Real>(256)));
        if (solve_threads == 512)
            BA_TRY(allow_dynamic_lds((const void*)k_wsolve<512>, (int)solve_lds()));
        else
            BA_TRY(allow_dynamic_lds((const void*)k_wsolve<256>, (int)solve_lds()));
        return BA_OK;
    }

    // dynamic LDS of k_wsolve: the packed factor of the largest window factored in LDS
    size_t solve_lds() const { return (size_t)work_nmax * (work_nmax + 1) / 2 * sizeof(double); }

    template <int M>
    int enqueue_iteration_m() {
        if (threads == 512)
            return solve_threads == 512 ? enqueue_iteration_nt<M, 512, 512>() : enqueue_iteration_nt<M, 512, 256>();
        return solve_threads == 512 ? enqueue_iteration_nt<M, 256, 512>() : enqueue_iteration_nt<M, 256, 256>();
    }

    int enqueue_iteration() {
        return model == kModelStereo ? enqueue_iteration_m<kModelStereo>() : enqueue_iteration_m<kModelSnavely>();
    }

    // Runs `iters` LM iterations of the windows in `slots` (dispatched in the given order); with
    // use_graph the iteration is captured once per work set and replayed.
    int run(const std::vector<int>& slots, int iters, bool use_graph = false, bool restart_clock = true) {
        if (slots.empty()) return BA_OK;
        BA_TRY(set_work(slots));
        BA_TRY(set_lds_limits());
        if (use_graph && !graph) {
            hipGraph_t g = nullptr;
            BA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            BA_TRY(enqueue_iteration());
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (e != hipSuccess) return fail(BA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            const hipError_t e2 = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e2 != hipSuccess) {
                graph = nullptr;
                return fail(BA_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e2));
            }
        }
        if (restart_clock) hipLaunchKernelGGL(k_wclock, dim3(1), dim3(64), 0, s, dclock.p);
        for (int it = 0; it < iters; ++it) {
            if (use_graph)
                BA_HIP(hipGraphLaunch(graph, s));
            else
                BA_TRY(enqueue_iteration());
            BA_HIP(hipGetLastError());
        }
        return BA_OK;
    }

    int reset(const std::vector<int>& slots) {
        if (slots.empty()) return BA_OK;
        BA_HIP(hipMemcpyAsync(dwork.p, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_window_reset, dim3((unsigned)slots.size()), dim3(256), 0, s, (const WinDesc*)ddesc.p,
                           arena.p, (const int*)dwork.p, (int)slots.size(), init_state());
        BA_HIP(hipGetLastError());
        return BA_OK;
    }

    // States of `slots` (synchronizes the stream).
    int states(const std::vector<int>& slots, std::vector<WinState>* out) {
        out->assign(slots.size(), WinState{});
        if (slots.empty()) return BA_OK;
        BA_HIP(hipMemcpyAsync(dwork.p, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_window_states, dim3(div_up((int64_t)slots.size(), 64)), dim3(64), 0, s,
                           (const WinDesc*)ddesc.p, (const char*)arena.p, (const int*)dwork.p, (int)slots.size(),
                           dstates.p);
        BA_HIP(hipMemcpyAsync(out->data(), dstates.p, slots.size() * sizeof(WinState), hipMemcpyDeviceToHost, s));
        BA_TRY(sync("window states"));
        return BA_OK;
    }

    int download(int slot, const std::vector<int>& pt_perm, double* cams, double* pts) {
        const WinDesc& D = hdesc[slot];
        std::vector<double> px(4 * (size_t)D.P);
        if (cams && D.K) BA_HIP(hipMemcpyAsync(cams, arena.p + D.o_cam_x, 72 * (size_t)D.K, hipMemcpyDeviceToHost, s));
        if (D.P) BA_HIP(hipMemcpyAsync(px.data(), arena.p + D.o_pt_x, 32 * (size_t)D.P, hipMemcpyDeviceToHost, s));
        BA_TRY(sync("window download"));
        if (pts)
            for (int ip = 0; ip < D.P; ++ip)
                for (int i = 0; i < 3; ++i) pts[3 * (size_t)pt_perm[ip] + i] = px[4 * (size_t)ip + i];
        return BA_OK;
    }

    // Times one LM iteration of the listed windows (ms), leaving their state as it was.
    int time_iteration(const std::vector<int>& slots, float* ms) {
        BA_TRY(reset(slots));
        BA_TRY(run(slots, 1));  // iteration 0 + first step (warm)
        BA_TRY(reset(slots));
        BA_HIP(hipEventRecord(ev0, s));
        BA_TRY(run(slots, 2));
        BA_HIP(hipEventRecord(ev1, s));
        BA_HIP(hipEventSynchronize(ev1));
        BA_HIP(hipEventElapsedTime(ms, ev0, ev1));
        return reset(slots);
    }
};

template <typename Real>
void fill_summary(const WinState& st, ba_summary* sm, bool g2o = false) {
    const double cscale = g2o ? 2.0 : 1.0;  // g2o reports chi2 = sum rho(s), without the 1/2
    sm->initial_cost = cscale * st.initial_cost;
    sm->final_cost = cscale * st.x_cost;
    sm->num_iterations = st.rows;
    sm->num_successful_steps = st.num_successful;
    sm->num_unsuccessful_steps = st.num_unsuccessful;
    sm->termination_type = st.termination;
    sm->termination_reason = st.reason;
    sm->termination_value = st.reason_value;
    sm->linear_solver_used = BA_SOLVER_DENSE_SCHUR;
    sm->total_linear_solver_iterations = st.iteration;
}

// ---------------------------------------------------------------------------------------------
// Batch
struct BatchBase {
    virtual ~BatchBase() = default;
    virtual int run(ba_summary* summaries) = 0;
    virtual int reset() = 0;
    virtual int download(int w, double* cams, double* pts) = 0;
    virtual int set_threads(int t) = 0;
    virtual void info(int* threads, int64_t* bytes) = 0;
    virtual int get_schedule(ba_batch_schedule* out) = 0;
    virtual int set_schedule(const ba_batch_schedule* sch) = 0;
};

// The windows of a batch are dealt to up to kBatchLanes lanes (largest work first, in turn), each an
// engine with its own arena, stream, task lists and graph: the lanes' launch sequences run
// concurrently, so one lane's low-occupancy phases (the per-window factorization, the small
// per-window kernels) overlap the others' wide ones.  Measured on 512 windows (the 512-window
// makespan, best of ten runs on one device): one lane 26.8 ms before the Schur-product and
// factorization changes; with the present kernels 2 lanes 20.0-20.3 ms, 3 lanes 19.9, 4 lanes
// 19.8, 6 lanes 22.1 and 8 lanes 21.6 (more lanes' Schur products contend for the L2).  A lane
// gets at least kBatchLaneMinWindows windows.  A window's result does not depend on its lane.
constexpr int kBatchLanes = 4;
constexpr int kBatchLaneMinWindows = 32;

template <typename Real>
struct Batch final : BatchBase {
    std::vector<std::unique_ptr<WindowEngine<Real>>> lanes;
    std::vector<std::vector<int>> lane_order;  // per lane: window slots in dispatch order
    std::vector<int> lane_of;                  // window -> lane (-1: single-problem solver)
    std::vector<WindowImage> imgs;
    std::vector<int> small;              // windows solved in the batched kernels
    std::vector<int> large;              // windows solved by the single-problem solver
    std::vector<std::unique_ptr<SolverBase>> large_solvers;
    std::vector<int> large_index;        // window -> index in large_solvers or -1
    std::vector<ba_summary> large_sum;
    ba_options opt{};
    ba_batch_schedule sched{};  // the schedule in use and the autotuner's timings
    std::vector<hipStream_t> lane_streams;  // from the device's pool (acquire_lane_streams)
    std::vector<bool> lane_pooled;

    ~Batch() override {
        for (auto& e : lanes) (void)hipStreamSynchronize(e->s);
        lanes.clear();
        if (!lane_streams.empty()) release_lane_streams(opt.device, lane_streams, lane_pooled);
    }

    // stereo: per window the full_slam stereo model (rig, sides), or NULL for the Snavely model.
    // The g2o rules are available with the stereo model (full_slam's solver); Snavely windows run the
    // Ceres rules (the g2o driver's rotation update is the single-problem solver's).
    int create(const ba_problem_view* views, int count, const ba_options& o, const StereoInput* stereo = nullptr,
               bool tune = true) {
        opt = o;
        BA_TRY(validate_options(&o));
        if (o.mode != BA_MODE_CERES && !stereo)
            return fail(BA_ERR_INVALID_ARGUMENT, "batched Snavely windows run the Ceres mode only");
        BA_TRY(use_device(o.device));
        imgs.resize(count);
        large_index.assign(count, -1);
        lane_of.assign(count, -1);
        for (int w = 0; w < count; ++w) {
            BA_TRY(validate_view(&views[w], true));
            const StereoInput* sw = stereo ? &stereo[w] : nullptr;
            if (9 * views[w].num_cameras > kWinMaxN) {
                large_index[w] = (int)large_solvers.size();
                ba_options lo = o;
                lo.linear_solver = BA_SOLVER_DENSE_SCHUR;
                std::unique_ptr<SolverBase> sv;
                BA_TRY(make_solver(&views[w], lo, &sv, sw));
                large_solvers.push_back(std::move(sv));
                large.push_back(w);
                continue;
            }
            BA_TRY(build_window_image<Real>(&views[w], o, &imgs[w], sw));
            small.push_back(w);
        }
        large_sum.assign(large_solvers.size(), ba_summary{});
        std::vector<int> order = small;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return imgs[a].cost > imgs[b].cost; });
        const int nl = std::max(1, std::min(kBatchLanes, (int)order.size() / kBatchLaneMinWindows));
        lane_order.assign(nl, {});
        for (size_t i = 0; i < order.size(); ++i) {
            lane_order[i % nl].push_back(order[i]);
            lane_of[order[i]] = (int)(i % nl);
        }
        BA_TRY(acquire_lane_streams(o.device, nl, &lane_streams, &lane_pooled));
        for (int l = 0; l < nl; ++l) {
            auto eng = std::make_unique<WindowEngine<Real>>();
            eng->model = stereo ? kModelStereo : kModelSnavely;
            int64_t total = 0;
            for (int w : lane_order[l]) total += imgs[w].total;
            BA_TRY(eng->init(o, total, count, lane_streams[l]));
            int64_t base = 0;
            for (int w : lane_order[l]) {
                BA_TRY(eng->place(imgs[w], w, base));
                base += imgs[w].total;
            }
            BA_HIP(hipStreamSynchronize(eng->s));
            lanes.push_back(std::move(eng));
        }
        // the host images are only needed for their permutations from here on
        for (int w : small) std::vector<char>().swap(imgs[w].bytes);
        if (tune) return autotune();
        sched = ba_batch_schedule{};
        apply(256, 256, 2, kWinMaxK);
        return BA_OK;
    }

    // Times schedules on the batch itself (all lanes running together, as run() runs them) and keeps
    // the fastest; the windows' results do not depend on the choice.  Two stages (the two kernels
    // they tune run one after the other): the factor's workgroup size and LDS limit with the default
    // Schur-product launch, then the Schur-product workgroup size and blocks per wave with the
    // chosen factor launch.  Each candidate is timed twice over two LM iterations from the initial
    // state (best of two).
    int autotune() {
        sched = ba_batch_schedule{};
        apply(256, 256, 2, kWinMaxK);
        if (small.empty()) return BA_OK;
        double best = 1e30;
        double warm = 0.0;
        BA_TRY(time_iteration(&warm));  // first launches of the kernels, not timed
        auto time_it = [&](double* ms) -> int {
            *ms = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                double m = 0.0;
                BA_TRY(time_iteration(&m));
                *ms = std::min(*ms, m / 2.0);
            }
            sched.candidates++;
            return BA_OK;
        };
        int best_tf = 256, best_lds = kWinMaxK;
        for (int tf : {256, 512})
            for (int lc : {kWinMaxK, 14, 10}) {
                apply(256, tf, 2, lc);
                double ms;
                BA_TRY(time_it(&ms));
                if (tf == 256 && lc == kWinMaxK) sched.default_iteration_ms = ms;
                if (ms < best) {
                    best = ms;
                    best_tf = tf;
                    best_lds = lc;
                }
            }
        int best_ts = 256, best_bpw = 2;
        for (int ts : {256, 512}) {
            if (schur_lds<Real>(ts) + 64 > lds_per_block()) continue;
            for (int bpw : {1, 2, 4}) {
                if (ts == 256 && bpw == 2) continue;  // timed in the first stage
                apply(ts, best_tf, bpw, best_lds);
                double ms;
                BA_TRY(time_it(&ms));
                if (ms < best) {
                    best = ms;
                    best_ts = ts;
                    best_bpw = bpw;
                }
            }
        }
        const int cand = sched.candidates;
        const double dflt = sched.default_iteration_ms;
        apply(best_ts, best_tf, best_bpw, best_lds);
        sched.candidates = cand;
        sched.default_iteration_ms = dflt;
        sched.iteration_ms = best;
        return BA_OK;
    }

    // Milliseconds of two LM iterations of every lane, launched together (host clock around the
    // launches and the waits); the windows are reset before and after.
    int time_iteration(double* ms) {
        const int nl = (int)lanes.size();
        for (int l = 0; l < nl; ++l) {
            BA_TRY(lanes[l]->reset(lane_order[l]));
            BA_TRY(lanes[l]->run(lane_order[l], 1));  // iteration 0 + first step (warm)
            BA_TRY(lanes[l]->reset(lane_order[l]));
        }
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->sync("autotune"));
        const auto t0 = std::chrono::steady_clock::now();
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->run(lane_order[l], 2));
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->sync("autotune"));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->reset(lane_order[l]));
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->sync("autotune"));
        return BA_OK;
    }

    static void apply_lane(WindowEngine<Real>& eng, int ts, int tf, int bpw, int lds_cameras) {
        eng.threads = ts;
        eng.solve_threads = tf;
        eng.blocks_per_wave = bpw;
        eng.lds_cameras = lds_cameras;
    }

    void apply(int ts, int tf, int bpw, int lds_cameras) {
        for (auto& e : lanes) apply_lane(*e, ts, tf, bpw, lds_cameras);
        sched.schur_threads = ts;
        sched.solve_threads = tf;
        sched.blocks_per_wave = bpw;
        sched.lds_cameras = lanes.empty() ? std::min(lds_cameras, kWinMaxK) : lanes[0]->lds_n() / 9;
        sched.lanes = (int)lanes.size();
    }

    int set_schedule(const ba_batch_schedule* sch) override {
        if (!sch) return autotune();
        const int ts = sch->schur_threads, tf = sch->solve_threads, bpw = sch->blocks_per_wave;
        if ((ts != 256 && ts != 512) || (tf != 256 && tf != 512))
            return fail(BA_ERR_INVALID_ARGUMENT, "workgroup sizes must be 256 or 512 (got %d, %d)", ts, tf);
        if (bpw != 1 && bpw != 2 && bpw != 4 && bpw != 8)
            return fail(BA_ERR_INVALID_ARGUMENT, "blocks_per_wave must be 1, 2, 4 or 8 (got %d)", bpw);
        if (sch->lds_cameras < 0) return fail(BA_ERR_INVALID_ARGUMENT, "lds_cameras < 0");
        if (schur_lds<Real>(ts) + 64 > lds_per_block())
            return fail(BA_ERR_CAPACITY, "the Schur staging of %d threads does not fit the LDS", ts);
        sched = ba_batch_schedule{};
        apply(ts, tf, bpw, sch->lds_cameras);
        return BA_OK;
    }

    int get_schedule(ba_batch_schedule* out) override {
        *out = sched;
        return BA_OK;
    }

    int set_threads(int t) override {
        if (t == 0) return autotune();
        ba_batch_schedule sc = sched;
        sc.schur_threads = sc.solve_threads = t;
        return set_schedule(&sc);
    }

    void info(int* threads, int64_t* bytes) override {
        if (threads) *threads = lanes.empty() ? 256 : lanes[0]->threads;
        if (bytes) {
            *bytes = 0;
            for (auto& e : lanes) *bytes += e->arena_cap;
        }
    }

    int run(ba_summary* summaries) override {
        BA_TRY(use_device(opt.device));
        const bool graphs = opt.launch_mode == 0;
        const int nl = (int)lanes.size();
        // one LM trial per launch sequence, the first after the initial linearization: under the
        // Ceres rules every window has decided its last trial after max_num_iterations sequences
        const int iters = std::max(1, opt.max_num_iterations);
        // the lanes run concurrently, each lane's launch sequences issued before the next lane's: the
        // lanes start a few graph launches apart, so their phases stay offset (issuing the lanes'
        // sequences in turn, which starts them together, measured 0.3 ms slower at 512 windows)
        for (int l = 0; l < nl; ++l) BA_TRY(lanes[l]->run(lane_order[l], iters, graphs));
        std::vector<std::vector<WinState>> st(nl);
        // a g2o iteration takes up to g2o_max_trials trials: continue four at a time while a window
        // of the lane is still running (under the Ceres rules this finds every window done)
        const int cap = (opt.mode == BA_MODE_G2O ? opt.max_num_iterations * std::max(opt.g2o_max_trials, 1)
                                                 : opt.max_num_iterations) + 1;
        std::vector<int> lane_iters(nl, iters);
        for (bool more_work = true; more_work;) {
            more_work = false;
            for (int l = 0; l < nl; ++l) {
                if (lane_iters[l] >= cap) continue;
                BA_TRY(lanes[l]->states(lane_order[l], &st[l]));
                if (std::all_of(st[l].begin(), st[l].end(), [](const WinState& x) { return x.done != 0; }))
                    continue;
                const int more = std::min(4, cap - lane_iters[l]);
                BA_TRY(lanes[l]->run(lane_order[l], more, graphs, false));
                lane_iters[l] += more;
                more_work = true;
            }
        }
        for (size_t i = 0; i < large.size(); ++i) BA_TRY(large_solvers[i]->run(&large_sum[i]));
        // finish times from the earliest lane start
        unsigned long long t0 = ~0ull;
        for (int l = 0; l < nl; ++l) {
            BA_TRY(lanes[l]->states(lane_order[l], &st[l]));
            unsigned long long tl = 0;
            BA_TRY(lanes[l]->clock_start(&tl));
            t0 = std::min(t0, tl);
        }
        if (summaries) {
            for (int l = 0; l < nl; ++l) {
                for (size_t i = 0; i < lane_order[l].size(); ++i) {
                    const WinState& ws = st[l][i];
                    ba_summary& sm = summaries[lane_order[l][i]];
                    ba_iteration* tr = sm.trace;
                    const int cap = sm.trace_capacity;
                    fill_summary<Real>(ws, &sm, opt.mode == BA_MODE_G2O);
                    sm.total_time_ms = ws.done && ws.t_done >= t0 ? (double)(ws.t_done - t0) / lanes[l]->clock_khz : 0.0;
                    sm.trace = tr;
                    sm.trace_capacity = cap;
                }
            }
            for (size_t i = 0; i < large.size(); ++i) {
                ba_summary& sm = summaries[large[i]];
                ba_iteration* tr = sm.trace;
                const int cap = sm.trace_capacity;
                sm = large_sum[i];
                sm.trace = tr;
                sm.trace_capacity = cap;
            }
        }
        return BA_OK;
    }

    int reset() override {
        for (size_t l = 0; l < lanes.size(); ++l) BA_TRY(lanes[l]->reset(lane_order[l]));
        for (auto& sv : large_solvers) BA_TRY(sv->reset());
        for (auto& e : lanes) BA_HIP(hipStreamSynchronize(e->s));
        return BA_OK;
    }

    int download(int w, double* cams, double* pts) override {
        if (w < 0 || w >= (int)imgs.size()) return fail(BA_ERR_INVALID_ARGUMENT, "window %d out of range", w);
        if (large_index[w] >= 0) return large_solvers[large_index[w]]->download(cams, pts);
        return lanes[lane_of[w]]->download(w, imgs[w].pt_perm, cams, pts);
    }
};

// Windows run in FP64 whatever options.precision asks (Real, the type of the per-observation records
// and of the Schur pair products, is double for both).  A window's reduced system has at most 288
// unknowns, where ba_solve factors in FP64 as well (kDenseF64MaxN): an FP32 factor would need CG
// against the FP64 system to reach the FP64 step (measured on the 512-window batch: 0.73 ms per LM
// iteration for the FP32 factor + CG against 0.32 ms for the FP64 factor), and FP32 records with
// the FP64 factor (Batch<float>) gave 22.0 ms against 23.5 ms for the 512-window batch but ended
// up to 4.7e-5 from the FP64 final cost on its worst windows.  Windows above 32 cameras go to the
// single-problem solver with the caller's precision.
int make_batch(const ba_problem_view* views, int count, const ba_options& o, std::unique_ptr<BatchBase>* out,
               const StereoInput* stereo = nullptr, bool tune = true) {
    auto b = std::make_unique<Batch<double>>();
    BA_TRY(b->create(views, count, o, stereo, tune));
    *out = std::move(b);
    return BA_OK;
}

// ---------------------------------------------------------------------------------------------
// Serving
struct ServerBase {
    virtual ~ServerBase() = default;
    virtual int submit(const ba_problem_view* v, int64_t tag) = 0;
    virtual int step(int* resident, int* queued) = 0;
    virtual int poll(ba_server_result* out, int cap, int* n) = 0;
    // Recorded from any thread, applied by the next step.
    virtual int cancel(int64_t tag) = 0;
    virtual int evict(int64_t tag) = 0;
    virtual double clock() const = 0;
    // One line of state for error messages: queue, ready images, resident windows, large problems.
    virtual std::string describe() = 0;
};

// Device waits of the server (and of its large-problem solvers) give up after this long, and the
// serving simulation reports an error with the server's state when no request has finished for
// kServerStallS seconds: a stuck state ends with a message instead of silence.
constexpr double kServerWaitS = 60.0;
constexpr double kServerStallS = 120.0;

struct Request {
    int64_t tag = 0;
    double submit = 0.0, admit = 0.0;
    int nc = 0, np = 0;
    std::vector<int> ci, pi;
    std::vector<double> obs, cams, pts;
    // device image, built by the server's builder threads after submit
    std::unique_ptr<WindowImage> img;
    int build_status = BA_OK;
    std::atomic<int> ready{0};
    // times the window left the device with its state (its image then holds the whole region:
    // parameters, LM state and working arrays, relocatable because every offset is window-local)
    int evictions = 0;
    ba_problem_view view() {
        ba_problem_view v{};
        v.num_cameras = nc;
        v.num_points = np;
        v.num_observations = (int)ci.size();
        v.camera_block_size = 9;
        v.camera_index = ci.data();
        v.point_index = pi.data();
        v.observations = obs.data();
        v.cameras = cams.data();
        v.points = pts.data();
        return v;
    }
};

struct Done {
    ba_server_result r{};
    std::vector<double> cams, pts;
};

template <typename Real>
struct Server final : ServerBase {
    ba_server_config cfg{};
    WindowEngine<Real> eng;
    std::chrono::steady_clock::time_point t0;
    std::map<int64_t, int64_t> free_ranges;  // offset -> size
    std::vector<int> free_slots;
    struct Resident {
        int slot = -1;
        int64_t off = 0, size = 0;
        std::unique_ptr<Request> req;
        WindowImage img;
    };
    std::vector<Resident> res;  // by slot
    std::vector<int> active;    // resident slots
    // submitted windows in arrival order; guarded by qmu (ba_server_submit may run on another
    // thread than ba_server_step)
    std::mutex qmu;
    std::deque<std::unique_ptr<Request>> queue;
    // cancellations and evictions recorded by cancel() / evict() (guarded by qmu), applied by step()
    enum { kOpCancel = 0, kOpEvict = 1 };
    std::vector<std::pair<int64_t, int>> ops;
    std::deque<Done> done;
    std::vector<Done> polled;   // keeps the last poll's parameter arrays alive
    // requests too large for a window slot: solved by the single-problem solver on worker threads
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Request>> big_queue;
    bool stopping = false;
    int big_inflight = 0;
    // builder threads: device images of submitted windows are built off the stepping thread
    std::vector<std::thread> builders;
    std::mutex bmu;
    std::condition_variable bcv;
    std::deque<Request*> build_queue;
    bool bstop = false;
    char* pinned = nullptr;  // staging of finished windows' parameters
    size_t pinned_cap = 0;

    ~Server() override {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv.notify_all();
        for (auto& t : workers)
            if (t.joinable()) t.join();
        {
            std::lock_guard<std::mutex> lk(bmu);
            bstop = true;
        }
        bcv.notify_all();
        for (auto& t : builders)
            if (t.joinable()) t.join();
        if (pinned) (void)hipHostFree(pinned);
    }

    void build_loop() {
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(bmu);
                bcv.wait(lk, [&] { return bstop || !build_queue.empty(); });
                if (bstop) return;
                r = build_queue.front();
                build_queue.pop_front();
            }
            auto img = std::make_unique<WindowImage>();
            ba_problem_view v = r->view();
            r->build_status = build_window_image<Real>(&v, cfg.options, img.get());
            r->img = std::move(img);
            r->ready.store(1, std::memory_order_release);
        }
    }

    double clock() const override {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    int create(const ba_server_config& c) {
        cfg = c;
        if (cfg.max_windows < 1) return fail(BA_ERR_INVALID_ARGUMENT, "max_windows < 1");
        if (cfg.arena_bytes < 1) return fail(BA_ERR_INVALID_ARGUMENT, "arena_bytes < 1");
        BA_TRY(validate_options(&cfg.options));
        if (cfg.options.mode != BA_MODE_CERES) return fail(BA_ERR_INVALID_ARGUMENT, "the server runs the Ceres mode only");
        BA_TRY(use_device(cfg.options.device));
        if (cfg.iterations_per_step < 1) cfg.iterations_per_step = 1;
        BA_TRY(eng.init(cfg.options, cfg.arena_bytes, cfg.max_windows));
        eng.wait_timeout_s = kServerWaitS;
        free_ranges[0] = eng.arena_cap;
        res.resize(cfg.max_windows);
        for (int i = cfg.max_windows - 1; i >= 0; --i) free_slots.push_back(i);
        t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < 2; ++i) workers.emplace_back([this] { big_loop(); });
        const int nb = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
        for (int i = 0; i < nb; ++i) builders.emplace_back([this] { build_loop(); });
        return BA_OK;
    }

    bool alloc(int64_t size, int64_t* off) {
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
            if (it->second >= size) {
                *off = it->first;
                const int64_t rest = it->second - size;
                const int64_t o2 = it->first + size;
                free_ranges.erase(it);
                if (rest > 0) free_ranges[o2] = rest;
                return true;
            }
        }
        return false;
    }

    void release(int64_t off, int64_t size) {
        auto it = free_ranges.emplace(off, size).first;
        auto next = std::next(it);
        if (next != free_ranges.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_ranges.erase(next);
        }
        if (it != free_ranges.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_ranges.erase(it);
            }
        }
    }

    int submit(const ba_problem_view* v, int64_t tag) override {
        BA_TRY(validate_view(v, true));
        auto r = std::make_unique<Request>();
        r->tag = tag;
        r->submit = clock();
        r->nc = v->num_cameras;
        r->np = v->num_points;
        r->ci.assign(v->camera_index, v->camera_index + v->num_observations);
        r->pi.assign(v->point_index, v->point_index + v->num_observations);
        r->obs.assign(v->observations, v->observations + 2 * (size_t)v->num_observations);
        r->cams.assign(v->cameras, v->cameras + 9 * (size_t)v->num_cameras);
        r->pts.assign(v->points, v->points + 3 * (size_t)v->num_points);
        if (9 * r->nc > kWinMaxN) {
            {
                std::lock_guard<std::mutex> lk(mu);
                big_queue.push_back(std::move(r));
                big_inflight++;
            }
            cv.notify_one();
            return BA_OK;
        }
        Request* rp = r.get();
        {
            std::lock_guard<std::mutex> lk(qmu);
            queue.push_back(std::move(r));
        }
        {
            std::lock_guard<std::mutex> lk(bmu);
            build_queue.push_back(rp);
        }
        bcv.notify_one();
        return BA_OK;
    }

    void big_loop() {
        for (;;) {
            std::unique_ptr<Request> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stopping || !big_queue.empty(); });
                if (stopping) return;
                r = std::move(big_queue.front());
                big_queue.pop_front();
            }
            r->admit = clock();
            ba_problem_view v = r->view();
            ba_options o = cfg.options;
            o.linear_solver = BA_SOLVER_AUTO;
            // stream launches, no graph capture: solvers are created, run and destroyed on two
            // worker threads at once, and a capture open on one thread while another frees device
            // or pinned memory (calls that synchronize the device) is an interleaving the server
            // avoids; the results are the same bit for bit
            o.launch_mode = 1;
            std::unique_ptr<SolverBase> sv;
            ba_summary sm{};
            Done d;
            d.r.tag = r->tag;
            d.r.evictions = r->evictions;
            d.r.submit_time_s = r->submit;
            d.r.admit_time_s = r->admit;
            int st = make_solver(&v, o, &sv);
            if (st == BA_OK) sv->set_wait_timeout(kServerWaitS);
            if (st == BA_OK) st = sv->run(&sm);
            if (st == BA_OK) st = sv->download(v.cameras, v.points);
            d.r.initial_cost = sm.initial_cost;
            d.r.final_cost = sm.final_cost;
            d.r.iterations = sm.num_iterations;
            d.r.termination_type = st == BA_OK ? sm.termination_type : BA_FAILURE;
            d.cams = std::move(r->cams);
            d.pts = std::move(r->pts);
            d.r.finish_time_s = clock();
            std::lock_guard<std::mutex> lk(mu);
            done.push_back(std::move(d));
            big_inflight--;
        }
    }

    // A request that can never be solved (its image failed to build, or it needs more device memory
    // than the whole arena) leaves the queue with a BA_FAILURE result instead of blocking the
    // requests behind it.
    void reject(std::unique_ptr<Request> r) {
        Done d;
        d.r.tag = r->tag;
        d.r.evictions = r->evictions;
        d.r.submit_time_s = r->submit;
        d.r.admit_time_s = d.r.finish_time_s = clock();
        d.r.termination_type = BA_FAILURE;
        d.cams = std::move(r->cams);
        d.pts = std::move(r->pts);
        std::lock_guard<std::mutex> lk(mu);
        done.push_back(std::move(d));
    }

    int cancel(int64_t tag) override {
        std::lock_guard<std::mutex> lk(qmu);
        ops.emplace_back(tag, (int)kOpCancel);
        return BA_OK;
    }

    int evict(int64_t tag) override {
        std::lock_guard<std::mutex> lk(qmu);
        ops.emplace_back(tag, (int)kOpEvict);
        return BA_OK;
    }

    // A cancelled request's result: BA_USER_FAILURE with the parameters it had (those of its last
    // accepted step when it had run; `st` is its state then, NULL if it never ran).
    void post_cancelled(std::unique_ptr<Request> r, const WinState* st) {
        Done d;
        d.r.tag = r->tag;
        d.r.evictions = r->evictions;
        d.r.submit_time_s = r->submit;
        d.r.admit_time_s = r->admit > 0.0 ? r->admit : clock();
        d.r.finish_time_s = clock();
        d.r.termination_type = BA_USER_FAILURE;
        d.r.initial_cost = d.r.final_cost = std::numeric_limits<double>::quiet_NaN();
        if (st) {
            d.r.initial_cost = st->initial_cost;
            d.r.final_cost = st->x_cost;
            d.r.iterations = st->rows;
        }
        d.cams = std::move(r->cams);
        d.pts = std::move(r->pts);
        std::lock_guard<std::mutex> lk(mu);
        done.push_back(std::move(d));
    }

    // Frees slot `sl` and its arena range; the resident's request is returned.
    std::unique_ptr<Request> vacate(int sl) {
        Resident& r = res[sl];
        std::unique_ptr<Request> q = std::move(r.req);
        release(r.off, r.size);
        r = Resident{};
        free_slots.push_back(sl);
        active.erase(std::find(active.begin(), active.end(), sl));
        return q;
    }

    // Cancels the resident window in slot `sl`: its current parameters and state are read back.
    int cancel_resident(int sl) {
        std::vector<WinState> st;
        BA_TRY(eng.states({sl}, &st));
        Request& q = *res[sl].req;
        BA_TRY(eng.download(sl, res[sl].img.pt_perm, q.cams.data(), q.pts.data()));
        post_cancelled(vacate(sl), &st[0]);
        return BA_OK;
    }

    // Evicts the resident window in slot `sl`: its whole region is copied to host memory and the
    // request rejoins the back of the queue, ready to be placed anywhere in the arena again.
    int evict_resident(int sl) {
        Resident& r = res[sl];
        auto img = std::make_unique<WindowImage>(std::move(r.img));
        img->bytes.resize((size_t)r.size);
        BA_HIP(hipMemcpyAsync(img->bytes.data(), eng.arena.p + r.off, (size_t)r.size, hipMemcpyDeviceToHost, eng.s));
        BA_TRY(eng.sync("evicted window"));
        std::unique_ptr<Request> q = vacate(sl);
        q->img = std::move(img);
        q->evictions++;
        std::lock_guard<std::mutex> lk(qmu);
        queue.push_back(std::move(q));
        return BA_OK;
    }

    // A queued request that had been evicted carries its parameters and state in its image.
    void cancel_queued(std::unique_ptr<Request> r) {
        if (r->evictions == 0 || !r->img) {
            post_cancelled(std::move(r), nullptr);
            return;
        }
        const WindowImage& img = *r->img;
        const WinDesc& D = img.desc;
        WinState st;
        std::memcpy(&st, img.bytes.data() + D.o_state, sizeof(WinState));
        std::memcpy(r->cams.data(), img.bytes.data() + D.o_cam_x, 72 * (size_t)D.K);
        const double* px = reinterpret_cast<const double*>(img.bytes.data() + D.o_pt_x);
        for (int ip = 0; ip < D.P; ++ip)
            for (int k = 0; k < 3; ++k) r->pts[3 * (size_t)img.pt_perm[ip] + k] = px[4 * (size_t)ip + k];
        post_cancelled(std::move(r), &st);
    }

    // Applies the recorded cancellations and evictions.  A cancellation of a request whose image
    // is still being built waits for a later step; tags that name nothing pending (finished,
    // unknown, a large problem already on the single-problem solver, an eviction of a window that
    // is not resident) are dropped.
    int apply_ops() {
        std::vector<std::pair<int64_t, int>> todo, later;
        {
            std::lock_guard<std::mutex> lk(qmu);
            todo.swap(ops);
        }
        for (const auto& op : todo) {
            const int64_t tag = op.first;
            const bool is_cancel = op.second == kOpCancel;
            const auto it = std::find_if(active.begin(), active.end(), [&](int sl) { return res[sl].req->tag == tag; });
            if (it != active.end()) {
                BA_TRY(is_cancel ? cancel_resident(*it) : evict_resident(*it));
                continue;
            }
            if (!is_cancel) continue;
            std::unique_ptr<Request> r;
            bool building = false;
            {
                std::lock_guard<std::mutex> lk(qmu);
                for (auto q = queue.begin(); q != queue.end(); ++q)
                    if ((*q)->tag == tag) {
                        if ((*q)->ready.load(std::memory_order_acquire)) {
                            r = std::move(*q);
                            queue.erase(q);
                        } else {
                            building = true;
                        }
                        break;
                    }
            }
            if (building) {
                later.push_back(op);
                continue;
            }
            if (r) {
                cancel_queued(std::move(r));
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                for (auto q = big_queue.begin(); q != big_queue.end(); ++q)
                    if ((*q)->tag == tag) {
                        r = std::move(*q);
                        big_queue.erase(q);
                        big_inflight--;
                        break;
                    }
            }
            if (r) post_cancelled(std::move(r), nullptr);
        }
        if (!later.empty()) {
            std::lock_guard<std::mutex> lk(qmu);
            ops.insert(ops.begin(), later.begin(), later.end());
        }
        return BA_OK;
    }

    int admit() {
        std::vector<int> placed;
        while (!free_slots.empty()) {
            Request* head = nullptr;
            {
                std::lock_guard<std::mutex> lk(qmu);
                if (queue.empty()) break;
                head = queue.front().get();
            }
            if (!head->ready.load(std::memory_order_acquire)) break;  // FIFO: wait for the head's image
            if (head->build_status != BA_OK || head->img->total > eng.arena_cap) {
                std::unique_ptr<Request> r;
                {
                    std::lock_guard<std::mutex> lk(qmu);
                    r = std::move(queue.front());
                    queue.pop_front();
                }
                reject(std::move(r));
                continue;
            }
            WindowImage& img = *head->img;
            int64_t off = 0;
            if (!alloc(img.total, &off)) break;  // wait for memory to free up (FIFO admission)
            const int slot = free_slots.back();
            free_slots.pop_back();
            Resident& r = res[slot];
            r.slot = slot;
            r.off = off;
            r.size = img.total;
            {
                std::lock_guard<std::mutex> lk(qmu);
                r.req = std::move(queue.front());
                queue.pop_front();
            }
            if (r.req->evictions == 0) r.req->admit = clock();  // the first admission
            r.img = std::move(*r.req->img);
            r.req->img.reset();
            BA_TRY(eng.place(r.img, slot, off));
            active.push_back(slot);
            placed.push_back(slot);
        }
        if (!placed.empty()) BA_TRY(eng.sync("window uploads"));
        for (int sl : placed) std::vector<char>().swap(res[sl].img.bytes);
        return BA_OK;
    }

    std::string describe() override {
        size_t nq = 0, nready = 0;
        {
            std::lock_guard<std::mutex> lk(qmu);
            nq = queue.size();
            for (const auto& r : queue) nready += r->ready.load(std::memory_order_acquire) ? 1 : 0;
        }
        size_t ndone = 0, nbig = 0;
        int inflight = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            ndone = done.size();
            nbig = big_queue.size();
            inflight = big_inflight;
        }
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "queued %zu (images ready %zu), resident %zu of %d slots, free arena ranges %zu, large problems "
                 "in flight %d (waiting %zu), results not yet polled %zu",
                 nq, nready, active.size(), cfg.max_windows, free_ranges.size(), inflight, nbig, ndone);
        return buf;
    }

    int step(int* resident, int* queued) override {
        BA_TRY(use_device(cfg.options.device));
        BA_TRY(apply_ops());
        BA_TRY(admit());
        if (!active.empty()) {
            std::vector<int> order = active;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return res[a].img.cost > res[b].img.cost; });
            BA_TRY(eng.run(order, cfg.iterations_per_step));
            std::vector<WinState> st;
            BA_TRY(eng.states(order, &st));
            // parameters of every finished window in one pass through pinned staging memory
            std::vector<size_t> stage_off(order.size(), 0);
            size_t need = 0;
            for (size_t i = 0; i < order.size(); ++i)
                if (st[i].done) {
                    const WinDesc& D = eng.hdesc[order[i]];
                    stage_off[i] = need;
                    need += 72 * (size_t)D.K + 32 * (size_t)D.P;
                }
            if (need > pinned_cap) {
                if (pinned) (void)hipHostFree(pinned);
                pinned = nullptr;
                pinned_cap = 0;
                BA_HIP(hipHostMalloc((void**)&pinned, need, hipHostMallocDefault));
                pinned_cap = need;
            }
            for (size_t i = 0; i < order.size(); ++i)
                if (st[i].done) {
                    const WinDesc& D = eng.hdesc[order[i]];
                    char* dst = pinned + stage_off[i];
                    if (D.K) BA_HIP(hipMemcpyAsync(dst, eng.arena.p + D.o_cam_x, 72 * (size_t)D.K, hipMemcpyDeviceToHost, eng.s));
                    if (D.P)
                        BA_HIP(hipMemcpyAsync(dst + 72 * (size_t)D.K, eng.arena.p + D.o_pt_x, 32 * (size_t)D.P,
                                              hipMemcpyDeviceToHost, eng.s));
                }
            if (need) BA_TRY(eng.sync("finished windows' parameters"));
            const double now = clock();
            for (size_t i = 0; i < order.size(); ++i) {
                if (!st[i].done) continue;
                const int sl = order[i];
                Resident& r = res[sl];
                const WinDesc& D = eng.hdesc[sl];
                Done d;
                d.r.tag = r.req->tag;
                d.r.evictions = r.req->evictions;
                d.r.submit_time_s = r.req->submit;
                d.r.admit_time_s = r.req->admit;
                d.r.finish_time_s = now;
                d.r.initial_cost = st[i].initial_cost;
                d.r.final_cost = st[i].x_cost;
                d.r.iterations = st[i].rows;
                d.r.termination_type = st[i].termination;
                d.cams = std::move(r.req->cams);
                d.pts = std::move(r.req->pts);
                const char* src = pinned + stage_off[i];
                std::memcpy(d.cams.data(), src, 72 * (size_t)D.K);
                const double* px = reinterpret_cast<const double*>(src + 72 * (size_t)D.K);
                for (int ip = 0; ip < D.P; ++ip)
                    for (int k = 0; k < 3; ++k) d.pts[3 * (size_t)r.img.pt_perm[ip] + k] = px[4 * (size_t)ip + k];
                {
                    std::lock_guard<std::mutex> lk(mu);
                    done.push_back(std::move(d));
                }
                release(r.off, r.size);
                r = Resident{};
                free_slots.push_back(sl);
                active.erase(std::find(active.begin(), active.end(), sl));
            }
        }
        if (resident) *resident = (int)active.size();
        if (queued) {
            int nq = 0;
            {
                std::lock_guard<std::mutex> lk(qmu);
                nq = (int)queue.size();
            }
            std::lock_guard<std::mutex> lk(mu);
            *queued = nq + big_inflight;
        }
        return BA_OK;
    }

    int poll(ba_server_result* out, int cap, int* n) override {
        std::lock_guard<std::mutex> lk(mu);
        polled.clear();
        int k = 0;
        while (k < cap && !done.empty()) {
            polled.push_back(std::move(done.front()));
            done.pop_front();
            ++k;
        }
        for (int i = 0; i < k; ++i) {
            Done& d = polled[i];
            out[i] = d.r;
            out[i].num_cameras = (int)(d.cams.size() / 9);
            out[i].num_points = (int)(d.pts.size() / 3);
            out[i].cameras = d.cams.data();
            out[i].points = d.pts.data();
        }
        if (n) *n = k;
        return BA_OK;
    }
};

// Resident windows run in FP64 as in make_batch; the large problems keep the caller's precision.
int make_server(const ba_server_config& c, std::unique_ptr<ServerBase>* out) {
    auto s = std::make_unique<Server<double>>();
    BA_TRY(s->create(c));
    *out = std::move(s);
    return BA_OK;
}

}  // namespace mba

struct ba_batch {
    std::unique_ptr<mba::BatchBase> impl;
    std::vector<ba_problem_view> views;
};

struct ba_server {
    std::unique_ptr<mba::ServerBase> impl;
};

using namespace mba;

extern "C" {

int ba_batch_create(const ba_problem_view* windows, int32_t count, const ba_options* options, ba_batch** out) {
    clear_error();
    if (!out || (count > 0 && !windows) || count < 0) return fail(BA_ERR_INVALID_ARGUMENT, "invalid arguments");
    ba_options o;
    if (options)
        o = *options;
    else
        ba_default_options(&o);
    auto h = std::make_unique<ba_batch>();
    BA_TRY(make_batch(windows, count, o, &h->impl));
    h->views.assign(windows, windows + count);
    *out = h.release();
    return BA_OK;
}

int ba_batch_run(ba_batch* b, ba_summary* summaries) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL batch");
    return b->impl->run(summaries);
}

int ba_batch_reset(ba_batch* b) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL batch");
    return b->impl->reset();
}

int ba_batch_download(ba_batch* b, int32_t w, double* cameras, double* points) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL batch");
    return b->impl->download(w, cameras, points);
}

int ba_batch_destroy(ba_batch* b) {
    delete b;
    return BA_OK;
}

int ba_batch_set_block(ba_batch* b, int32_t threads) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL batch");
    return b->impl->set_threads(threads);
}

int ba_batch_get_schedule(ba_batch* b, ba_batch_schedule* out) {
    clear_error();
    if (!b || !out) return fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    return b->impl->get_schedule(out);
}

int ba_batch_set_schedule(ba_batch* b, const ba_batch_schedule* schedule) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "batch is NULL");
    return b->impl->set_schedule(schedule);
}

int ba_batch_info(ba_batch* b, int32_t* threads, int64_t* device_bytes) {
    clear_error();
    if (!b) return fail(BA_ERR_INVALID_ARGUMENT, "NULL batch");
    int t = 0;
    int64_t bytes = 0;
    b->impl->info(&t, &bytes);
    if (threads) *threads = t;
    if (device_bytes) *device_bytes = bytes;
    return BA_OK;
}

int ba_batch_solve(const ba_problem_view* windows, int32_t count, const ba_options* options, ba_summary* summaries) {
    ba_batch* b = nullptr;
    BA_TRY(ba_batch_create(windows, count, options, &b));
    std::unique_ptr<ba_batch> guard(b);
    BA_TRY(b->impl->run(summaries));
    for (int w = 0; w < count; ++w) BA_TRY(b->impl->download(w, windows[w].cameras, windows[w].points));
    return BA_OK;
}

int ba_stereo_batch_solve(const ba_stereo_view* windows, int32_t count, const ba_options* options,
                          ba_summary* summaries, uint8_t* const* outliers, double* chi2_thresholds) {
    clear_error();
    if (count < 0 || (count > 0 && !windows)) return fail(BA_ERR_INVALID_ARGUMENT, "invalid window array");
    if (count == 0) return BA_OK;
    ba_options o;
    if (options)
        o = *options;
    else
        ba_stereo_default_options(&o);
    BA_TRY(validate_options(&o));
    std::vector<StereoProblem> sp(count);
    std::vector<ba_problem_view> views(count);
    std::vector<StereoInput> in(count);
    for (int w = 0; w < count; ++w) {
        const int st = stereo_prepare(&windows[w], &sp[w]);
        if (st != BA_OK) return fail(st, "window %d: %s", w, ba_last_error_string());
        views[w] = sp[w].view;
        in[w] = StereoInput{&sp[w].rig, windows[w].side};
    }
    // one solve per layout: the default schedule (autotuning would time more iterations than the solve)
    std::unique_ptr<BatchBase> b;
    BA_TRY(make_batch(views.data(), count, o, &b, in.data(), false));
    BA_TRY(b->run(summaries));
    for (int w = 0; w < count; ++w) {
        BA_TRY(b->download(w, sp[w].cams.data(), windows[w].points));
        stereo_finish(&windows[w], sp[w], o.loss_scale, outliers ? outliers[w] : nullptr,
                      chi2_thresholds ? &chi2_thresholds[w] : nullptr);
    }
    return BA_OK;
}

int ba_server_create(const ba_server_config* config, ba_server** out) {
    clear_error();
    if (!config || !out) return fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    auto h = std::make_unique<ba_server>();
    BA_TRY(make_server(*config, &h->impl));
    *out = h.release();
    return BA_OK;
}

int ba_server_submit(ba_server* s, const ba_problem_view* window, int64_t tag) {
    clear_error();
    if (!s) return fail(BA_ERR_INVALID_ARGUMENT, "NULL server");
    return s->impl->submit(window, tag);
}

int ba_server_step(ba_server* s, int32_t* resident, int32_t* queued) {
    clear_error();
    if (!s) return fail(BA_ERR_INVALID_ARGUMENT, "NULL server");
    int r = 0, q = 0;
    BA_TRY(s->impl->step(&r, &q));
    if (resident) *resident = r;
    if (queued) *queued = q;
    return BA_OK;
}

int ba_server_cancel(ba_server* s, int64_t tag) {
    clear_error();
    if (!s) return fail(BA_ERR_INVALID_ARGUMENT, "NULL server");
    return s->impl->cancel(tag);
}

int ba_server_evict(ba_server* s, int64_t tag) {
    clear_error();
    if (!s) return fail(BA_ERR_INVALID_ARGUMENT, "NULL server");
    return s->impl->evict(tag);
}

int ba_server_poll(ba_server* s, ba_server_result* out, int32_t capacity, int32_t* count) {
    clear_error();
    if (!s || (!out && capacity > 0)) return fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    int n = 0;
    BA_TRY(s->impl->poll(out, capacity, &n));
    if (count) *count = n;
    return BA_OK;
}

double ba_server_clock(ba_server* s) { return s ? s->impl->clock() : 0.0; }

int ba_server_destroy(ba_server* s) {
    delete s;
    return BA_OK;
}

int ba_server_simulate(const ba_server_config* config, const ba_problem_view* pool, int32_t pool_count, double rate_hz,
                       int32_t count, uint64_t seed, ba_server_result* results) {
    clear_error();
    if (!config || !pool || pool_count < 1 || count < 0 || !(rate_hz > 0.0) || (count > 0 && !results))
        return fail(BA_ERR_INVALID_ARGUMENT, "invalid arguments");
    ba_server* sv = nullptr;
    BA_TRY(ba_server_create(config, &sv));
    std::unique_ptr<ba_server> guard(sv);
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_hz);
    std::vector<double> arrival(count);
    double t = 0.0;
    for (int i = 0; i < count; ++i) {
        t += gap(rng);
        arrival[i] = t;
    }
    const double start = sv->impl->clock();
    int next = 0, finished = 0;
    double last_progress = 0.0;
    std::vector<ba_server_result> buf(64);
    while (finished < count) {
        const double now = sv->impl->clock() - start;
        if (next == finished) last_progress = now;  // nothing outstanding: waiting for arrivals
        if (now - last_progress > kServerStallS) {
            const std::string state = sv->impl->describe();
            return fail(BA_ERR_TIMEOUT,
                        "serving simulation: no request finished for %.0f s (submitted %d of %d, finished %d); %s",
                        now - last_progress, next, count, finished, state.c_str());
        }
        while (next < count && arrival[next] <= now) {
            BA_TRY(sv->impl->submit(&pool[next % pool_count], next));
            ++next;
        }
        int resident = 0, queued = 0;
        BA_TRY(sv->impl->step(&resident, &queued));
        int n = 0;
        do {
            BA_TRY(sv->impl->poll(buf.data(), (int)buf.size(), &n));
            for (int i = 0; i < n; ++i) {
                ba_server_result r = buf[i];
                const int64_t id = r.tag;
                // latency is measured from the scheduled arrival time
                r.submit_time_s = arrival[id];
                r.admit_time_s -= start;
                r.finish_time_s -= start;
                r.cameras = nullptr;
                r.points = nullptr;
                results[id] = r;
                ++finished;
                last_progress = sv->impl->clock() - start;
            }
        } while (n == (int)buf.size());
        if (resident == 0) {
            // idle (or only requests on the single-problem solver): wait for the next arrival
            const double wait = next < count ? arrival[next] - (sv->impl->clock() - start) : 1e-4;
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
        }
    }
    return BA_OK;
}

}  // extern "C"
