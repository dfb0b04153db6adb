// eval.hip -- per-observation Snavely residual + Jacobian on the GPU (the parity entry point that
// replaces one AutoDiffCostFunction<SnavelyReprojectionError,2,9,3> evaluation per residual block,
// utils/camera/SnavelyReprojectionError.h:42-92), and the FP64 cost 1/2 sum rho(|r|^2).
//
// Layout: observations in caller order; cameras AoS (9 doubles) as in BALProblem.  One thread per
// observation, camera prep computed once per camera by k_camera_prep_aos.
#include <map>
#include <mutex>

#include "common.h"
#include "snavely.h"

namespace mba {

template <typename Real>
__global__ void k_camera_prep_aos(const double* __restrict__ cams, int nc, CameraPrep<Real>* __restrict__ prep) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double cam[9];
    for (int i = 0; i < 9; ++i) cam[i] = cams[9 * (size_t)c + i];
    CameraPrep<Real> P;
    camera_prep<Real>(cam, P);
    prep[c] = P;
}

template <typename Real>
__global__ void k_eval_obs(int no, const int* __restrict__ ci, const int* __restrict__ pi,
                           const double* __restrict__ obs, const double* __restrict__ pts,
                           const CameraPrep<Real>* __restrict__ prep, double* __restrict__ r_out,
                           double* __restrict__ jc_out, double* __restrict__ jp_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= no) return;
    const CameraPrep<Real> P = prep[ci[k]];
    const int p = pi[k];
    const double X[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
    Real r[2], Jc[18], Jp[6];
    snavely_residual_jacobian<Real>(P, X, (Real)obs[2 * (size_t)k], (Real)obs[2 * (size_t)k + 1], r, Jc, Jp);
    r_out[2 * (size_t)k] = (double)r[0];
    r_out[2 * (size_t)k + 1] = (double)r[1];
    if (jc_out)
        for (int i = 0; i < 18; ++i) jc_out[18 * (size_t)k + i] = (double)Jc[i];
    if (jp_out)
        for (int i = 0; i < 6; ++i) jp_out[6 * (size_t)k + i] = (double)Jp[i];
}

// Per-block partial sums of 1/2 rho(|r|^2) in FP64, reduced in a fixed order by k_sum_partials.
__global__ void k_cost_obs(int no, const int* __restrict__ ci, const int* __restrict__ pi,
                           const double* __restrict__ obs, const double* __restrict__ pts,
                           const CameraPrep<double>* __restrict__ prep, int loss, double a,
                           double* __restrict__ partial) {
    __shared__ double red[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double c = 0.0;
    if (k < no) {
        const CameraPrep<double> P = prep[ci[k]];
        const int p = pi[k];
        const double X[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
        double r[2];
        snavely_residual<double>(P, X, obs[2 * (size_t)k], obs[2 * (size_t)k + 1], r);
        c = half_rho(r[0] * r[0] + r[1] * r[1], loss, a);
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// k_cost_obs for indices that the host has not validated (their arrays live in device memory):
// an observation whose camera or point index is out of range contributes NaN, so the cost is NaN.
__global__ void k_cost_obs_checked(int no, int nc, int np, const int* __restrict__ ci, const int* __restrict__ pi,
                                   const double* __restrict__ obs, const double* __restrict__ pts,
                                   const CameraPrep<double>* __restrict__ prep, int loss, double a,
                                   double* __restrict__ partial) {
    __shared__ double red[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double c = 0.0;
    if (k < no) {
        const int cam = ci[k], p = pi[k];
        if (cam < 0 || cam >= nc || p < 0 || p >= np) {
            c = __builtin_nan("");
        } else {
            const CameraPrep<double> P = prep[cam];
            const double X[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
            double r[2];
            snavely_residual<double>(P, X, obs[2 * (size_t)k], obs[2 * (size_t)k + 1], r);
            c = half_rho(r[0] * r[0] + r[1] * r[1], loss, a);
        }
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------------------------------------
// Allocation-free cost evaluation (ba_eval_cost_noalloc): the kernels of ba_eval_cost with their
// work space (camera preps, per-block partial sums) in a file-scope device array instead of memory
// the library allocates.  Each GPU holds its own copy until the process ends; host code uses the
// address hipGetSymbolAddress returns for the current device, never the array's name.
constexpr int kNoAllocMaxCameras = 4096;
constexpr int kNoAllocMaxObservations = 1 << 24;
constexpr size_t kNoAllocPrepBytes = sizeof(CameraPrep<double>) * (size_t)kNoAllocMaxCameras;
constexpr size_t kNoAllocPartials = (size_t)kNoAllocMaxObservations / 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocPrepBytes + kNoAllocPartials * sizeof(double);
static_assert(kNoAllocPrepBytes % 256 == 0, "partial sums start 256-byte aligned");

__device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per device.
int noalloc_scratch(char** out) {
    static std::mutex mu;
    static std::map<int, char*> addr;  // device -> address of its copy of scratch
    int dev = 0;
    BA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(BA_ERR_HIP, "scratch buffer address on device %d: %s", dev, hipGetErrorString(e));
        it = addr.emplace(dev, static_cast<char*>(ptr)).first;
    }
    *out = it->second;
    return BA_OK;
}

struct EvalUpload {
    DevBuf<int> ci, pi;
    DevBuf<double> obs, cams, pts;
    int upload(const ba_problem_view* v, hipStream_t s) {
        BA_TRY(ci.upload(v->camera_index, (size_t)v->num_observations, s));
        BA_TRY(pi.upload(v->point_index, (size_t)v->num_observations, s));
        BA_TRY(obs.upload(v->observations, 2 * (size_t)v->num_observations, s));
        BA_TRY(cams.upload(v->cameras, 9 * (size_t)v->num_cameras, s));
        BA_TRY(pts.upload(v->points, 3 * (size_t)v->num_points, s));
        return BA_OK;
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) (void)hipStreamDestroy(s);
    }
};

template <typename Real>
int eval_impl(const ba_problem_view* v, double* residuals, double* jc, double* jp, hipStream_t s) {
    EvalUpload u;
    BA_TRY(u.upload(v, s));
    const int no = v->num_observations, nc = v->num_cameras;
    DevBuf<CameraPrep<Real>> prep;
    DevBuf<double> r, Jc, Jp;
    BA_TRY(prep.alloc(nc));
    BA_TRY(r.alloc(2 * (size_t)no));
    if (jc) BA_TRY(Jc.alloc(18 * (size_t)no));
    if (jp) BA_TRY(Jp.alloc(6 * (size_t)no));
    hipLaunchKernelGGL(k_camera_prep_aos<Real>, dim3(div_up(nc, 64)), dim3(64), 0, s, u.cams.p, nc, prep.p);
    BA_HIP(hipGetLastError());
    if (no > 0) {
        hipLaunchKernelGGL(k_eval_obs<Real>, dim3(div_up(no, 256)), dim3(256), 0, s, no, u.ci.p, u.pi.p, u.obs.p,
                           u.pts.p, prep.p, r.p, jc ? Jc.p : nullptr, jp ? Jp.p : nullptr);
        BA_HIP(hipGetLastError());
    }
    BA_HIP(hipMemcpyAsync(residuals, r.p, 2 * (size_t)no * sizeof(double), hipMemcpyDeviceToHost, s));
    if (jc) BA_HIP(hipMemcpyAsync(jc, Jc.p, 18 * (size_t)no * sizeof(double), hipMemcpyDeviceToHost, s));
    if (jp) BA_HIP(hipMemcpyAsync(jp, Jp.p, 6 * (size_t)no * sizeof(double), hipMemcpyDeviceToHost, s));
    BA_HIP(hipStreamSynchronize(s));
    return BA_OK;
}

}  // namespace
}  // namespace mba

using namespace mba;

extern "C" int ba_eval_residual_jacobian(const ba_problem_view* v, int32_t precision, double* residuals,
                                         double* jac_camera, double* jac_point) {
    clear_error();
    BA_TRY(validate_view(v, true));
    if (!residuals) return fail(BA_ERR_INVALID_ARGUMENT, "residuals output is NULL");
    if (precision != BA_PRECISION_FP64 && precision != BA_PRECISION_MIXED)
        return fail(BA_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
    BA_TRY(use_device(0));
    StreamGuard g;
    BA_HIP(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
    if (precision == BA_PRECISION_FP64) return eval_impl<double>(v, residuals, jac_camera, jac_point, g.s);
    return eval_impl<float>(v, residuals, jac_camera, jac_point, g.s);
}

extern "C" int ba_eval_cost(const ba_problem_view* v, int32_t loss, double loss_scale, double* cost) {
    clear_error();
    BA_TRY(validate_view(v, true));
    if (!cost) return fail(BA_ERR_INVALID_ARGUMENT, "cost output is NULL");
    BA_TRY(use_device(0));
    StreamGuard g;
    BA_HIP(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
    EvalUpload u;
    BA_TRY(u.upload(v, g.s));
    const int no = v->num_observations, nc = v->num_cameras;
    DevBuf<CameraPrep<double>> prep;
    BA_TRY(prep.alloc(nc));
    const int nb = no > 0 ? div_up(no, 256) : 1;
    DevBuf<double> partial, out;
    BA_TRY(partial.alloc(nb));
    BA_TRY(out.alloc(1));
    hipLaunchKernelGGL(k_camera_prep_aos<double>, dim3(div_up(nc, 64)), dim3(64), 0, g.s, u.cams.p, nc, prep.p);
    hipLaunchKernelGGL(k_cost_obs, dim3(nb), dim3(256), 0, g.s, no, u.ci.p, u.pi.p, u.obs.p, u.pts.p, prep.p, loss,
                       loss_scale, partial.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.s, partial.p, nb, out.p);
    BA_HIP(hipGetLastError());
    BA_HIP(hipMemcpyAsync(cost, out.p, sizeof(double), hipMemcpyDeviceToHost, g.s));
    BA_HIP(hipStreamSynchronize(g.s));
    return BA_OK;
}

extern "C" int ba_eval_cost_noalloc(const ba_problem_view* v, int32_t loss, double loss_scale, double* cost,
                                    void* stream) {
    clear_error();
    if (!v || !cost) return fail(BA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (v->camera_block_size != 9) return fail(BA_ERR_INVALID_ARGUMENT, "camera_block_size must be 9");
    const int no = v->num_observations, nc = v->num_cameras, np = v->num_points;
    if (no < 0 || nc < 0 || np < 0) return fail(BA_ERR_INVALID_ARGUMENT, "negative problem size");
    if (nc > kNoAllocMaxCameras || no > kNoAllocMaxObservations)
        return fail(BA_ERR_CAPACITY, "ba_eval_cost_noalloc holds at most %d cameras and %d observations (got %d, %d)",
                    kNoAllocMaxCameras, kNoAllocMaxObservations, nc, no);
    if (no > 0 && (!v->camera_index || !v->point_index || !v->observations || !v->points))
        return fail(BA_ERR_INVALID_ARGUMENT, "observation or point arrays are NULL");
    if (nc > 0 && !v->cameras) return fail(BA_ERR_INVALID_ARGUMENT, "camera array is NULL");
    char* base = nullptr;
    BA_TRY(noalloc_scratch(&base));
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto* prep = reinterpret_cast<CameraPrep<double>*>(base);
    auto* partial = reinterpret_cast<double*>(base + kNoAllocPrepBytes);
    const int nb = no > 0 ? div_up(no, 256) : 1;
    if (nc > 0)
        hipLaunchKernelGGL(k_camera_prep_aos<double>, dim3(div_up(nc, 64)), dim3(64), 0, s, v->cameras, nc, prep);
    hipLaunchKernelGGL(k_cost_obs_checked, dim3(nb), dim3(256), 0, s, no, nc, np, v->camera_index, v->point_index,
                       v->observations, v->points, (const CameraPrep<double>*)prep, loss, loss_scale, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double*)partial, nb, cost);
    BA_HIP(hipGetLastError());
    return BA_OK;
}
