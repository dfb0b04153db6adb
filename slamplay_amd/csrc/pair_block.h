// pair_block.h -- one 9x9 block of a reduced camera system from its observation pairs, by one wave.
//
// S_ab = [a == b] D_a^2 - sum_pairs J~_c,i' G_ij J~_c,j over the pairs (i, j) of observations of
// one point in cameras a and b, with G_ij = H_i J~_p,j' (H = J~_p V*^-1) for i != j and, for an
// observation with itself, -(I - G)_i (its share of U* folded in).  Used by the batched windows
// (records of 32 values, I - G formed from H and J~_p in working precision) and by the dense
// single-problem solver (records of 36 values with I - G formed in FP64 and stored).
//
// Per stage of 16 pairs the wave copies the record chunks the products read into its LDS staging
// area with 4-element vector loads (lanes cover the chunks of consecutive pairs; the next stage's
// records and the pair indices two stages ahead are in flight, requested in that order so that
// neither wait covers the other), every lane storing (spare slots write padding rows); then each
// row of 16 lanes takes one column group (0-2, 3-4, 5-6, 7-8) of all nine rows for the stage's 16
// pairs, lane q of the row pair q, accumulating in registers (FP64); the 16 partial sums of an entry
// are combined by a DPP butterfly inside the row.  The block's entries are written to S (row-major,
// leading dimension ld, both triangles).  Every entry's sum runs in a fixed order (pair q, q + 16,
// ... per lane, then the butterfly), whatever wave or launch takes the block.
#pragma once

#include "lm_kernels.h"

namespace mba {

constexpr int kPairsPerStage = 16;
// Rows of each half (i records, j records) of a wave's staging area: the stage's pairs plus the rows
// the copy's spare slots write to, so that every lane stores.
constexpr int kStageRows = kPairsPerStage + 4;

// Two consecutive record elements, read as one LDS access.
template <typename Real>
struct alignas(2 * sizeof(Real)) R2v {
    Real x, y;
};

// Record layouts.  Both keep J~_c by columns (the two residual rows of column c at 2c, 2c + 1).
//   kRecWindow: [J~_c 18][H 6][J~_p 6][2 unused]; i records read chunks 0-5, j records 0-4 and 6-7.
//   kRecDense:  [J~_c 18][J~_p 6][H 6][I - G: g00 g01 g11][3 unused]; i records read chunks 0-4 and
//               6-8 (J~_c, H, I - G), j records 0-5 (J~_c, J~_p).
// The staging pitch (34 / 38 values) puts the 16 records a row of lanes reads at once on distinct
// 4-bank groups.
enum { kRecWindow = 0, kRecDense = 1 };
template <int LAYOUT>
struct RecLayout;
template <>
struct RecLayout<kRecWindow> {
    static constexpr int stride = 32, pitch = 34, ci = 6, cp = 13, h = 18, jp = 24, ic = 0;
    static constexpr bool stored_complement = false;
    // chunk of the record for slot element e (e < ci: i record, else j record)
    __device__ static int chunk(int e) { return e < ci ? e : (e - ci < 5 ? e - ci : e - ci + 1); }
};
template <>
struct RecLayout<kRecDense> {
    static constexpr int stride = 36, pitch = 38, ci = 8, cp = 14, h = 24, jp = 18, ic = 30;
    static constexpr bool stored_complement = true;
    __device__ static int chunk(int e) { return e < ci ? (e < 5 ? e : e + 1) : e - ci; }
};

// LDS bytes of a wave's staging area.
template <typename Real, int LAYOUT>
constexpr int pair_stage_bytes() {
    return 2 * kStageRows * RecLayout<LAYOUT>::pitch * (int)sizeof(Real);
}

// One 64-lane round of the stage copy: slot s = 64 r + lane of the stage's m pairs (cp record chunks
// of four values per pair).  Slots past the stage's pairs load a chunk of its last pair (always a
// valid address, so the load needs no branch); their rows are not read.
template <typename Real, int LAYOUT>
__device__ __forceinline__ R4<Real> pair_load_round(const Real* __restrict__ js, int2 pr, int m, int r) {
    using RL = RecLayout<LAYOUT>;
    const int s = 64 * r + (threadIdx.x & 63);
    const int k = min(s / RL::cp, m - 1), e = s - RL::cp * (s / RL::cp);
    const int oi = __shfl(pr.x, k, 64), oj = __shfl(pr.y, k, 64);
    return reinterpret_cast<const R4<Real>*>(js + RL::stride * (size_t)(e < RL::ci ? oi : oj))[RL::chunk(e)];
}

template <typename Real, int LAYOUT>
__device__ __forceinline__ void pair_store_round(Real* st_i, Real* st_j, int r, const R4<Real>& v) {
    using RL = RecLayout<LAYOUT>;
    static_assert(4 * 64 <= kStageRows * RL::cp, "spare slots stay inside the staging area");
    static_assert(kPairsPerStage * RL::cp <= 4 * 64, "four copy rounds per stage");
    const int s = 64 * r + (threadIdx.x & 63);
    const int k = s / RL::cp, e = s - RL::cp * k;
    *reinterpret_cast<R4<Real>*>((e < RL::ci ? st_i : st_j) + RL::pitch * k + 4 * RL::chunk(e)) = v;
}

// Block bk = (camera a, camera b, first pair, end pair) of the pair list `pairs` (observation
// indices (i, j) into the records js) into S; dc2 holds D^2 (added on the diagonal of a == b).
template <typename Real, int LAYOUT>
__device__ __forceinline__ void pair_block(const int4 bk, const int2* __restrict__ pairs, const Real* __restrict__ js,
                                           const double* __restrict__ dc2, double* S, int ld, Real* stage) {
    using RL = RecLayout<LAYOUT>;
    constexpr int PS = kPairsPerStage;
    static_assert(PS == 16, "one pair per lane of a 16-lane row");
    const int lane = threadIdx.x & 63, q = lane & 15, h = lane >> 4;
    const int c0 = h == 0 ? 0 : 2 * h + 1;  // first column of the lane's group
    const int nc = h == 0 ? 3 : 2;          // columns in the group (a third slot repeats column 8 or c0 + 1)
    Real* st_i = stage;
    Real* st_j = stage + kStageRows * RL::pitch;
    double acc[9][3];
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[r][k] = 0.0;
    const int npairs = bk.w - bk.z;
    const int nst = (npairs + PS - 1) / PS;
    int2 pr = make_int2(0, 0), pr_next = make_int2(0, 0);
    if (lane < min(PS, npairs)) pr = pairs[bk.z + lane];
    if (nst > 1 && lane < min(PS, npairs - PS)) pr_next = pairs[bk.z + PS + lane];
    int m_ld = min(PS, npairs);
    R4<Real> b0 = pair_load_round<Real, LAYOUT>(js, pr, m_ld, 0), b1 = pair_load_round<Real, LAYOUT>(js, pr, m_ld, 1),
             b2 = pair_load_round<Real, LAYOUT>(js, pr, m_ld, 2), b3 = pair_load_round<Real, LAYOUT>(js, pr, m_ld, 3);
    for (int stg = 0; stg < nst; ++stg) {
        const int base = bk.z + stg * PS;
        const int m = min(PS, bk.w - base);
        pair_store_round<Real, LAYOUT>(st_i, st_j, 0, b0);
        pair_store_round<Real, LAYOUT>(st_i, st_j, 1, b1);
        pair_store_round<Real, LAYOUT>(st_i, st_j, 2, b2);
        pair_store_round<Real, LAYOUT>(st_i, st_j, 3, b3);
        const int pi = __shfl(pr.x, q, 64), pj = __shfl(pr.y, q, 64);
        // the pair indices two stages ahead are requested before the next stage's records and
        // taken over after this stage's products, so that neither waits for the records (a wave's
        // loads complete in order)
        int2 pr_far = make_int2(0, 0);
        if (stg + 1 < nst) {
            if (stg + 2 < nst && lane < min(PS, bk.w - base - 2 * PS)) pr_far = pairs[base + 2 * PS + lane];
            m_ld = min(PS, bk.w - base - PS);
            b0 = pair_load_round<Real, LAYOUT>(js, pr_next, m_ld, 0);
            b1 = pair_load_round<Real, LAYOUT>(js, pr_next, m_ld, 1);
            b2 = pair_load_round<Real, LAYOUT>(js, pr_next, m_ld, 2);
            b3 = pair_load_round<Real, LAYOUT>(js, pr_next, m_ld, 3);
        }
        wave_sync_lds();
        if (q < m) {
            using R2 = R2v<Real>;
            const Real* Ri = st_i + RL::pitch * q;
            const Real* Rj = st_j + RL::pitch * q;
            double G00, G01, G10, G11;
            if (RL::stored_complement && pi == pj) {
                // an observation with itself: -(I - G) as formed in FP64 when the record was written
                G00 = -(double)Ri[RL::ic];
                G01 = G10 = -(double)Ri[RL::ic + 1];
                G11 = -(double)Ri[RL::ic + 2];
            } else {
                const R2 h0 = *reinterpret_cast<const R2*>(Ri + RL::h), h1 = *reinterpret_cast<const R2*>(Ri + RL::h + 2),
                         h2 = *reinterpret_cast<const R2*>(Ri + RL::h + 4);
                const R2 p0 = *reinterpret_cast<const R2*>(Rj + RL::jp), p1 = *reinterpret_cast<const R2*>(Rj + RL::jp + 2),
                         p2 = *reinterpret_cast<const R2*>(Rj + RL::jp + 4);
                // H_i = [h0.x h0.y h1.x; h1.y h2.x h2.y], J~_p,j = [p0.x p0.y p1.x; p1.y p2.x p2.y]
                G00 = (double)(h0.x * p0.x + h0.y * p0.y + h1.x * p1.x);
                G01 = (double)(h0.x * p1.y + h0.y * p2.x + h1.x * p2.y);
                G10 = (double)(h1.y * p0.x + h2.x * p0.y + h2.y * p1.x);
                G11 = (double)(h1.y * p1.y + h2.x * p2.x + h2.y * p2.y);
                if (!RL::stored_complement) {
                    // an observation with itself carries its share of U*: J~c'(I - G)J~c, negated
                    // because the accumulated sum is subtracted
                    const double self = pi == pj ? 1.0 : 0.0;
                    G00 -= self;
                    G11 -= self;
                }
            }
            double t0[3], t1[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int c = min(c0 + k, 8);
                const R2 jc = *reinterpret_cast<const R2*>(Rj + 2 * c);
                t0[k] = G00 * (double)jc.x + G01 * (double)jc.y;
                t1[k] = G10 * (double)jc.x + G11 * (double)jc.y;
            }
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const R2 a = *reinterpret_cast<const R2*>(Ri + 2 * r);
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[r][k] = fma((double)a.y, t1[k], fma((double)a.x, t0[k], acc[r][k]));
            }
        }
        wave_sync_lds();
        pr = pr_next;
        pr_next = pr_far;
    }
    const int ca = bk.x, cb = bk.y;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double v = row16_sum(acc[r][k]);
            const int c = c0 + k;
            if (q == 0 && k < nc) {
                const double s = (ca == cb && r == c ? dc2[9 * ca + r] : 0.0) - v;
                S[(size_t)(9 * ca + r) * ld + 9 * cb + c] = s;
                if (ca != cb) S[(size_t)(9 * cb + c) * ld + 9 * ca + r] = s;
            }
        }
    }
}

}  // namespace mba
