// preprocess.hip -- BALProblem::Normalize and ::Perturb on the parameters a solver holds on the
// device (utils/datasets/BALProblem.cpp:39-76,230-336).
//
// Built with -ffp-contract=off so every product and sum rounds as the host code's does; with that,
// the medians, the L1 distances, the scale and the point transform are bit-identical to the host
// Normalize, and the perturbation draws the same normals as the host Perturb (equal to the last ulp
// of log()).
#include <cmath>
#include <mutex>
#include <vector>

#include "preprocess.h"

namespace mba {
namespace {

// ---------------------------------------------------------------------------------------------
// glibc rand() (random_r.c, TYPE_3: degree 31, separation 3).  srand(seed) sets r_0 = seed (1 for
// 0) and r_i = 16807 r_{i-1} mod (2^31 - 1) for i < 31 by Schrage's method in 32-bit arithmetic,
// then r_31..r_33 = r_0..r_2 and r_i = r_{i-31} + r_{i-3} (mod 2^32) onwards; srand discards 310
// outputs, so rand() call k returns r_{344+k} >> 1.  The state (r_{i-31}, ..., r_{i-1}) evolves by
// a fixed 31x31 matrix over Z/2^32, which is what makes jump-ahead possible.
constexpr int kDeg = 31;
constexpr int kChunk = 31 * 8;   // draws per generator thread
constexpr int kJumps = 24;       // generator threads below 2^24
constexpr int kPairsPerThread = 4;
constexpr int kPairBlock = 256;
constexpr int kPairsPerBlock = kPairsPerThread * kPairBlock;
constexpr int kSelBlock = 256;

struct RandState {
    uint32_t s[kDeg];
};

// state before rand() call 0: (r_313, ..., r_343)
RandState glibc_seed_state(uint32_t seed) {
    std::vector<uint32_t> r(344);
    if (seed == 0) seed = 1;
    int32_t word = (int32_t)seed;
    r[0] = (uint32_t)word;
    for (int i = 1; i < kDeg; ++i) {
        const int64_t hi = word / 127773, lo = word % 127773;
        int64_t next = 16807 * lo - 2836 * hi;
        if (next < 0) next += 2147483647;
        word = (int32_t)next;
        r[i] = (uint32_t)word;
    }
    for (int i = kDeg; i < 34; ++i) r[i] = r[i - kDeg];
    for (int i = 34; i < 344; ++i) r[i] = r[i - 31] + r[i - 3];
    RandState st;
    for (int j = 0; j < kDeg; ++j) st.s[j] = r[313 + j];
    return st;
}

using Mat31 = std::vector<uint32_t>;  // row-major 31x31

Mat31 mat_mul(const Mat31& a, const Mat31& b) {
    Mat31 c(kDeg * kDeg, 0u);
    for (int i = 0; i < kDeg; ++i)
        for (int k = 0; k < kDeg; ++k) {
            const uint32_t aik = a[i * kDeg + k];
            if (!aik) continue;
            for (int j = 0; j < kDeg; ++j) c[i * kDeg + j] += aik * b[k * kDeg + j];
        }
    return c;
}

// Q_j = A^(kChunk 2^j) for j < kJumps, A the one-draw transition.
const std::vector<uint32_t>& jump_tables() {
    static std::vector<uint32_t> tables;
    static std::once_flag once;
    std::call_once(once, [] {
        Mat31 a(kDeg * kDeg, 0u), p(kDeg * kDeg, 0u);
        for (int j = 0; j + 1 < kDeg; ++j) a[j * kDeg + j + 1] = 1u;  // shift
        a[(kDeg - 1) * kDeg + 0] = 1u;                                // r_{i-31}
        a[(kDeg - 1) * kDeg + 28] = 1u;                               // r_{i-3}
        for (int j = 0; j < kDeg; ++j) p[j * kDeg + j] = 1u;
        Mat31 base = a;
        for (int e = kChunk; e; e >>= 1) {  // p = A^kChunk
            if (e & 1) p = mat_mul(p, base);
            base = mat_mul(base, base);
        }
        tables.reserve((size_t)kJumps * kDeg * kDeg);
        for (int j = 0; j < kJumps; ++j) {
            tables.insert(tables.end(), p.begin(), p.end());
            p = mat_mul(p, p);
        }
    });
    return tables;
}

// One thread per chunk of kChunk draws: jump from the seed state to the chunk start, then run the
// recurrence 31 draws at a time with the state in registers (s[j] += s[j+28] for the first three,
// whose r_{i-3} is still the previous round's, s[j] += s[j-3] for the rest).  Each round's 64 x 31
// draws of the workgroup go through LDS so that lanes store consecutive draws of one chunk.
__global__ void __launch_bounds__(64) k_rand_generate(RandState start, const uint32_t* __restrict__ tables,
                                                      int n_chunks, int64_t n_draws, uint32_t* __restrict__ out) {
    __shared__ uint32_t tile[64][kDeg + 2];  // odd row stride: the per-thread row writes hit distinct banks
    const int b = blockIdx.x * 64 + threadIdx.x;
    const int lane_c = threadIdx.x / kDeg, lane_j = threadIdx.x < 2 * kDeg ? threadIdx.x - lane_c * kDeg : kDeg;
    uint32_t s[kDeg];
#pragma unroll
    for (int j = 0; j < kDeg; ++j) s[j] = start.s[j];
    for (int q = 0; q < kJumps && b < n_chunks && (b >> q); ++q) {
        if (!((b >> q) & 1)) continue;
        const uint32_t* Q = tables + (size_t)q * kDeg * kDeg;
        uint32_t t[kDeg];
#pragma unroll
        for (int i = 0; i < kDeg; ++i) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < kDeg; ++k) acc += Q[i * kDeg + k] * s[k];
            t[i] = acc;
        }
#pragma unroll
        for (int j = 0; j < kDeg; ++j) s[j] = t[j];
    }
    for (int round = 0; round < kChunk / kDeg; ++round) {
#pragma unroll
        for (int j = 0; j < 3; ++j) s[j] += s[j + 28];
#pragma unroll
        for (int j = 3; j < kDeg; ++j) s[j] += s[j - 3];
#pragma unroll
        for (int j = 0; j < kDeg; ++j) tile[threadIdx.x][j] = s[j] >> 1;
        __syncthreads();
        // lanes 0-30 and 31-61 store the 31 draws of two chunks per instruction
        if (lane_j < kDeg)
            for (int c = lane_c; c < 64; c += 2) {
                const int64_t bb = (int64_t)blockIdx.x * 64 + c;
                const int64_t o = bb * kChunk + (int64_t)round * kDeg + lane_j;
                if (bb < n_chunks && o < n_draws) out[o] = tile[c][lane_j];
            }
        __syncthreads();
    }
}

// BALProblem.cpp:39-55: x = 2 rand()/RAND_MAX - 1 for both coordinates, rejected when w >= 1 or 0.
__device__ __forceinline__ bool polar_pair(const uint32_t* __restrict__ draws, int64_t j, double* x1, double* w) {
    const double a = 2.0 * ((double)draws[2 * j] / 2147483647.0) - 1.0;
    const double c = 2.0 * ((double)draws[2 * j + 1] / 2147483647.0) - 1.0;
    const double ww = a * a + c * c;
    *x1 = a;
    *w = ww;
    return !(ww >= 1.0 || ww == 0.0);
}

// exclusive prefix sum over a 256-thread block; returns the block total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int wave_off = 0, all = 0;
    for (int k = 0; k < kPairBlock / 64; ++k) {
        const int c = lds[k];
        if (k < wave) wave_off += c;
        all += c;
    }
    __syncthreads();
    *total = all;
    return wave_off + x - v;
}

__global__ void __launch_bounds__(kPairBlock) k_pairs_count(const uint32_t* __restrict__ draws, int64_t n_pairs,
                                                            int* __restrict__ counts) {
    __shared__ int lds[kPairBlock / 64];
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock + (int64_t)threadIdx.x * kPairsPerThread;
    int c = 0;
    for (int k = 0; k < kPairsPerThread; ++k) {
        double x1, w;
        if (p0 + k < n_pairs && polar_pair(draws, p0 + k, &x1, &w)) ++c;
    }
    int total;
    block_exclusive_scan(c, lds, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// one block: counts -> exclusive offsets; total accepted pairs into *total
__global__ void __launch_bounds__(kPairBlock) k_pairs_scan(int* __restrict__ counts, int n, int* __restrict__ total) {
    __shared__ int lds[kPairBlock / 64];
    int carry = 0;
    for (int base = 0; base < n; base += kPairBlock) {
        const int i = base + (int)threadIdx.x;
        const int v = i < n ? counts[i] : 0;
        int tile;
        const int ex = block_exclusive_scan(v, lds, &tile);
        if (i < n) counts[i] = carry + ex;
        carry += tile;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kPairBlock) k_pairs_emit(const uint32_t* __restrict__ draws, int64_t n_pairs,
                                                           const int* __restrict__ offsets, int n_normals,
                                                           double* __restrict__ normals) {
    __shared__ int lds[kPairBlock / 64];
    const int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock + (int64_t)threadIdx.x * kPairsPerThread;
    double x1[kPairsPerThread], w[kPairsPerThread];
    bool ok[kPairsPerThread];
    int c = 0;
#pragma unroll
    for (int k = 0; k < kPairsPerThread; ++k) {
        ok[k] = p0 + k < n_pairs && polar_pair(draws, p0 + k, &x1[k], &w[k]);
        c += ok[k] ? 1 : 0;
    }
    int total;
    int idx = offsets[blockIdx.x] + block_exclusive_scan(c, lds, &total);
#pragma unroll
    for (int k = 0; k < kPairsPerThread; ++k)
        if (ok[k]) {
            if (idx < n_normals) normals[idx] = x1[k] * sqrt((-2.0 * log(w[k])) / w[k]);
            ++idx;
        }
}

// rotation.h:133-198 (AngleAxisRotatePoint), in the expression order of the host BALProblem code
__device__ __forceinline__ void rotate_point(const double aa[3], const double pt[3], double out[3]) {
    const double theta2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
    if (theta2 > 2.220446049250313080847e-16) {
        const double theta = sqrt(theta2);
        const double c = cos(theta), s = sin(theta);
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

// BALProblem.cpp:230-246: c = -R' t, through R(-aa)
__device__ __forceinline__ void camera_center(const double aa[3], const double t[3], double center[3]) {
    const double inv[3] = {-aa[0], -aa[1], -aa[2]};
    rotate_point(inv, t, center);
    for (int i = 0; i < 3; ++i) center[i] *= -1.0;
}

// BALProblem.cpp:248-261: t = -R c
__device__ __forceinline__ void center_translation(const double aa[3], const double center[3], double t[3]) {
    rotate_point(aa, center, t);
    for (int i = 0; i < 3; ++i) t[i] *= -1.0;
}

// BALProblem.cpp:305-336, points: X += n * sigma with the normals of the caller's point order
__global__ void k_perturb_points(double* __restrict__ pt, const int* __restrict__ pt_orig, int np,
                                 const double* __restrict__ normals, double sigma) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const int64_t o = 3 * (int64_t)pt_orig[ip];
    for (int j = 0; j < 3; ++j) pt[4 * (int64_t)ip + j] += normals[o + j] * sigma;
}

// cameras: angle-axis += n * rs with the centre held, t recomputed, then t += n * ts
__global__ void k_perturb_cams(double* __restrict__ cam, int nc, const double* __restrict__ normals, int64_t first,
                               double rs, double ts) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int per = (rs > 0.0 ? 3 : 0) + (ts > 0.0 ? 3 : 0);
    int64_t k = first + (int64_t)per * c;
    double aa[3], t[3], center[3];
    for (int i = 0; i < 3; ++i) {
        aa[i] = cam[(size_t)i * nc + c];
        t[i] = cam[(size_t)(3 + i) * nc + c];
    }
    camera_center(aa, t, center);
    if (rs > 0.0) {
        for (int i = 0; i < 3; ++i) aa[i] += normals[k + i] * rs;
        k += 3;
    }
    center_translation(aa, center, t);
    if (ts > 0.0)
        for (int i = 0; i < 3; ++i) t[i] += normals[k + i] * ts;
    for (int i = 0; i < 3; ++i) {
        cam[(size_t)i * nc + c] = aa[i];
        cam[(size_t)(3 + i) * nc + c] = t[i];
    }
}

// ---------------------------------------------------------------------------------------------
// Upper median (nth_element at n/2, BALProblem.cpp:71-76) by radix selection over the IEEE bit
// patterns mapped to order-preserving keys, 8 bits per pass from the top.  Selection k (0..2: a
// point axis, 3: the L1 distances) keeps (prefix, mask, rank) in sel[3k..3k+2].
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

__global__ void __launch_bounds__(256) k_sel_init(unsigned long long* __restrict__ sel, unsigned long long rank,
                                                  unsigned int* __restrict__ hist) {
    const int k = threadIdx.x;
    if (k < 4) {
        sel[3 * k] = 0;
        sel[3 * k + 1] = 0;
        sel[3 * k + 2] = rank;
    }
    for (int i = k; i < 4 * 256; i += 256) hist[i] = 0;
}

__global__ void __launch_bounds__(kSelBlock) k_sel_hist(const double* __restrict__ pt, const double* __restrict__ dist,
                                                        int np, const unsigned long long* __restrict__ sel,
                                                        int first, int shift, unsigned int* __restrict__ hist) {
    __shared__ unsigned int h[256];
    const int k = first + (int)blockIdx.y;
    const unsigned long long prefix = sel[3 * k], mask = sel[3 * k + 1];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int i = blockIdx.x * kSelBlock + threadIdx.x; i < np; i += gridDim.x * kSelBlock) {
        const double x = k < 3 ? pt[4 * (int64_t)i + k] : dist[i];
        const unsigned long long key = order_key(x);
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255], 1u);
    }
    __syncthreads();
    // integer counts: the sum does not depend on the order the blocks add in
    if (h[threadIdx.x]) atomicAdd(&hist[256 * k + threadIdx.x], h[threadIdx.x]);
}

// digit counts scanned; the thread whose bin holds the rank records it, and the counts are cleared
// for the next pass
__global__ void __launch_bounds__(256) k_sel_pick(unsigned long long* __restrict__ sel, unsigned int* __restrict__ hist,
                                                  int first, int shift) {
    __shared__ unsigned int wsum[4];
    const int k = first + (int)blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const unsigned int tot = hist[256 * k + t];
    hist[256 * k + t] = 0;
    unsigned int x = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned int off = 0;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    const unsigned long long incl = off + x, excl = incl - tot;
    const unsigned long long r = sel[3 * k + 2];
    __syncthreads();  // every thread has read the rank before it is replaced
    if (excl <= r && r < incl) {
        sel[3 * k] |= (unsigned long long)t << shift;
        sel[3 * k + 1] |= 255ull << shift;
        sel[3 * k + 2] = r - excl;
    }
}

// BALProblem.cpp:274-281: L1 distance of each point to the per-axis medians
__global__ void k_norm_dist(const double* __restrict__ pt, int np, const unsigned long long* __restrict__ sel,
                            double* __restrict__ dist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double* p = pt + 4 * (int64_t)i;
    dist[i] = fabs(p[0] - key_value(sel[0])) + fabs(p[1] - key_value(sel[3])) + fabs(p[2] - key_value(sel[6]));
}

// BALProblem.cpp:283-302: X <- s (X - m), camera centres likewise, s = 100 / MAD
__global__ void k_norm_apply(double* __restrict__ pt, int np, double* __restrict__ cam, int nc,
                             const unsigned long long* __restrict__ sel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double m[3] = {key_value(sel[0]), key_value(sel[3]), key_value(sel[6])};
    const double scale = 100.0 / key_value(sel[9]);
    if (i < np)
        for (int j = 0; j < 3; ++j) pt[4 * (int64_t)i + j] = scale * (pt[4 * (int64_t)i + j] - m[j]);
    if (i < nc) {
        double aa[3], t[3], center[3];
        for (int j = 0; j < 3; ++j) {
            aa[j] = cam[(size_t)j * nc + i];
            t[j] = cam[(size_t)(3 + j) * nc + i];
        }
        camera_center(aa, t, center);
        for (int j = 0; j < 3; ++j) center[j] = scale * (center[j] - m[j]);
        center_translation(aa, center, t);
        for (int j = 0; j < 3; ++j) cam[(size_t)(3 + j) * nc + i] = t[j];
    }
}

template <typename T>
int ensure(DevBuf<T>& b, size_t n) {
    return b.n >= n ? BA_OK : b.alloc(n);
}

}  // namespace

int draw_normals(int64_t n_normals, uint32_t seed, PrepScratch* w, hipStream_t s);

int device_normalize(const DeviceParams& p, PrepScratch* w, hipStream_t s) {
    if (p.np == 0) return BA_OK;  // BALProblem.cpp:263-303 leaves an empty problem as it is
    const int np = p.np;
    const int nblocks = std::max(1, std::min(256, div_up(np, kSelBlock * 8)));
    BA_TRY(ensure(w->sel_state, 12));
    BA_TRY(ensure(w->hist, (size_t)4 * 256));
    BA_TRY(ensure(w->dist, (size_t)np));
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, s, w->sel_state.p, (unsigned long long)(np / 2), w->hist.p);
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        hipLaunchKernelGGL(k_sel_hist, dim3(nblocks, 3), dim3(kSelBlock), 0, s, p.pt, w->dist.p, np, w->sel_state.p, 0,
                           shift, w->hist.p);
        hipLaunchKernelGGL(k_sel_pick, dim3(3), dim3(256), 0, s, w->sel_state.p, w->hist.p, 0, shift);
    }
    hipLaunchKernelGGL(k_norm_dist, dim3(div_up(np, 256)), dim3(256), 0, s, p.pt, np, w->sel_state.p, w->dist.p);
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        hipLaunchKernelGGL(k_sel_hist, dim3(nblocks, 1), dim3(kSelBlock), 0, s, p.pt, w->dist.p, np, w->sel_state.p, 3,
                           shift, w->hist.p);
        hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(256), 0, s, w->sel_state.p, w->hist.p, 3, shift);
    }
    hipLaunchKernelGGL(k_norm_apply, dim3(div_up(std::max(np, p.nc), 256)), dim3(256), 0, s, p.pt, np, p.cam, p.nc,
                       w->sel_state.p);
    BA_HIP(hipGetLastError());
    return BA_OK;
}

int device_perturb(const DeviceParams& p, uint32_t seed, double rs, double ts, double ps, PrepScratch* w,
                   hipStream_t s) {
    const int64_t n_pt = ps > 0.0 ? 3 * (int64_t)p.np : 0;
    const int64_t n_normals = n_pt + (int64_t)p.nc * ((rs > 0.0 ? 3 : 0) + (ts > 0.0 ? 3 : 0));
    if (n_normals > (int64_t)1 << 30) return fail(BA_ERR_CAPACITY, "perturbation needs %lld normals (limit 2^30)",
                                                   (long long)n_normals);
    if (n_normals > 0) BA_TRY(draw_normals(n_normals, seed, w, s));
    if (ps > 0.0 && p.np > 0)
        hipLaunchKernelGGL(k_perturb_points, dim3(div_up(p.np, 256)), dim3(256), 0, s, p.pt, p.pt_orig, p.np,
                           (const double*)w->normals.p, ps);
    // every camera goes through centre and back, as in the reference, also with both sigmas 0
    if (p.nc > 0)
        hipLaunchKernelGGL(k_perturb_cams, dim3(div_up(p.nc, 64)), dim3(64), 0, s, p.cam, p.nc,
                           (const double*)w->normals.p, n_pt, rs, ts);
    BA_HIP(hipGetLastError());
    return BA_OK;
}

// normals 0..n_normals-1 of the stream glibc's rand() gives after srand(seed), into w->normals
int draw_normals(int64_t n_normals, uint32_t seed, PrepScratch* w, hipStream_t s) {
    const std::vector<uint32_t>& tables = jump_tables();
    if (!w->table_chunks) {
        BA_TRY(w->tables.upload(tables.data(), tables.size(), s));
        w->table_chunks = 1;
    }
    const RandState start = glibc_seed_state(seed);
    // pairs to draw: the acceptance rate is pi/4, so 1.3 per normal leaves a margin of many
    // standard deviations; a shortfall (read back below) draws again with more
    int64_t n_pairs = n_normals + n_normals * 3 / 10 + 4096;
    BA_TRY(ensure(w->total, 1));
    for (;;) {
        const int64_t n_draws = 2 * n_pairs;
        const int n_chunks = div_up(n_draws, kChunk);
        if (n_chunks >= (1 << kJumps)) return fail(BA_ERR_CAPACITY, "random stream too long");
        const int nb = div_up(n_pairs, kPairsPerBlock);
        BA_TRY(ensure(w->draws, (size_t)n_draws));
        BA_TRY(ensure(w->counts, (size_t)nb));
        BA_TRY(ensure(w->normals, (size_t)n_normals));
        hipLaunchKernelGGL(k_rand_generate, dim3(div_up(n_chunks, 64)), dim3(64), 0, s, start, w->tables.p, n_chunks,
                           n_draws, w->draws.p);
        hipLaunchKernelGGL(k_pairs_count, dim3(nb), dim3(kPairBlock), 0, s, w->draws.p, n_pairs, (int*)w->counts.p);
        hipLaunchKernelGGL(k_pairs_scan, dim3(1), dim3(kPairBlock), 0, s, (int*)w->counts.p, nb, w->total.p);
        hipLaunchKernelGGL(k_pairs_emit, dim3(nb), dim3(kPairBlock), 0, s, w->draws.p, n_pairs,
                           (const int*)w->counts.p, (int)n_normals, w->normals.p);
        BA_HIP(hipGetLastError());
        int got = 0;
        BA_HIP(hipMemcpyAsync(&got, w->total.p, sizeof(int), hipMemcpyDeviceToHost, s));
        BA_HIP(hipStreamSynchronize(s));
        if (got >= n_normals) break;
        n_pairs = n_pairs + n_pairs / 2;
    }
    return BA_OK;
}

}  // namespace mba
