// window.h -- device layout of batched local-BA windows.
//
// Every window lives in one contiguous region of a device arena; WinDesc holds the absolute
// byte offsets of its arrays.  Within a window:
//   * observations in camera order (cam_off, co_pt local point index, co_uv64);
//   * points in SELL-64 chunks (chunk_off / chunk_len), slot_of[j] = slot of observation j; the
//     slot records pv are three double4 arrays of n_slots entries (component-major);
//   * the 9x9 blocks (a <= b) of the reduced camera system with their observation pairs
//     (i, j) (camera(i) = a, camera(j) = b, same point), built once on the host and listed with
//     the largest pair count first;
//   * per-camera and per-chunk partial sums (cpart, ppart) through which the phase kernels hand
//     results to the per-window kernels, which combine them in a fixed order.
#pragma once

#include <cstdint>

#include "lm_kernels.h"

namespace mba {

struct WinState {
    double radius, decrease_factor, x_cost, initial_cost, gmax;
    int iteration, consecutive_invalid, done, termination;
    int need_lin, first_lin, num_successful, num_unsuccessful;
    int rows;       // trace rows of the equivalent single-problem solve (iteration 0 included)
    int invalid;    // this iteration's damped point blocks were not all positive definite
    int have_step;  // this iteration produced a step
    int g2o_iter, g2o_trials;  // g2o mode: iterations done, rejected trials of the current one
    int reason;     // ba_termination_reason of the test that ended the solve
    unsigned long long t_done;  // device wall clock when the window finished
    double reason_value;        // the quantity that test compared
};

struct WinDesc {
    int K, P, No, n_chunks, n_slots, n_blocks, n_pairs, pad;
    int64_t o_cam_off, o_co_pt, o_co_uv64, o_slot_of, o_chunk_len, o_chunk_off, o_blk, o_pairs;
    int64_t o_cam_x, o_cam_x0, o_cam_c, o_pt_x, o_pt_x0, o_pt_c;
    int64_t o_prep, o_prep_c, o_udiag, o_gc, o_sc, o_dc2, o_bc, o_yc, o_cpart, o_ppart;
    int64_t o_V, o_gp, o_sp, o_Vinv, o_qp, o_yp, o_pv, o_js, o_S, o_L;
    int64_t o_side, o_rig;  // stereo model: camera side per camera-order observation, the rig
    int64_t o_state;
};

// Per-observation record (the FP64 residual and Jacobians are recomputed where needed):
//   js (FP64, 32 values): scaled J~_c column by column (18: the two rows of column c at 2c,
//       2c + 1), H = J~_p V*^-1 (6, row-major), J~_p (6, row-major), unused (2) -- the input of
//       the reduced camera system, rewritten whenever the damping changes; a pair product reads
//       the first 24 values of one record and values 0-17, 24-29 of the other.
constexpr int kWinJs = 32;
// Partial sums per camera (cpart) and per point chunk (ppart).
enum WinCamPart { WC_COST = 0, WC_GMAX, WC_MODEL, WC_CAND, WC_DX2, WC_X2, WC_COUNT = 8 };
enum WinPtPart { WP_GMAX = 0, WP_DX2, WP_X2, WP_MODEL, WP_COUNT = 4 };

template <typename T>
__host__ __device__ __forceinline__ T* at(char* base, int64_t off) {
    return reinterpret_cast<T*>(base + off);
}

}  // namespace mba
