// window.h -- device layout of batched local-BA windows (one workgroup per window).
//
// Every window lives in one contiguous region of a device arena; WinDesc holds the absolute
// byte offsets of its arrays.  Within a window:
//   * observations in camera order (cam_off, co_pt local point index, co_uv);
//   * points in SELL-64 chunks (chunk_off / chunk_len), slot_of[j] = slot of observation j;
//   * the 9x9 blocks (a <= b) of the reduced camera system with their observation pairs
//     (i, j) (camera(i) = a, camera(j) = b, same point), built once on the host.
#pragma once

#include <cstdint>

#include "lm_kernels.h"

namespace mba {

struct WinState {
    double radius, decrease_factor, x_cost, initial_cost, gmax;
    int iteration, consecutive_invalid, done, termination;
    int need_lin, first_lin, num_successful, num_unsuccessful;
    int rows, pad;  // trace rows of the equivalent single-problem solve (iteration 0 included)
};

struct WinDesc {
    int K, P, No, n_chunks, n_slots, n_blocks, n_pairs, pad;
    int64_t o_cam_off, o_co_pt, o_co_uv, o_co_uv64, o_slot_of, o_chunk_len, o_chunk_off, o_blk, o_pairs;
    int64_t o_cam_x, o_cam_x0, o_cam_c, o_pt_x, o_pt_x0, o_pt_c;
    int64_t o_prep, o_prep_c, o_udiag, o_gc, o_sc, o_dc2, o_bc, o_yc;
    int64_t o_V, o_gp, o_sp, o_Vinv, o_qp, o_yp, o_pv, o_u, o_js, o_S64, o_state;
};

constexpr int kWinJs = 36;  // per-observation record: Jc~ 18, Jp~ 6, Hp 6, I - G 3, r~ 2, pad 1

template <typename T>
__host__ __device__ __forceinline__ T* at(char* base, int64_t off) {
    return reinterpret_cast<T*>(base + off);
}

}  // namespace mba
