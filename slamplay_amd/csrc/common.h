// common.h -- error reporting, device buffers and launch helpers shared by the engine's sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mba/ba_c_api.h"

namespace mba {

// Records the message returned by ba_last_error_string() on this thread and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define BA_HIP(expr)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return ::mba::fail(e_ == hipErrorOutOfMemory ? BA_ERR_OUT_OF_MEMORY : BA_ERR_HIP,               \
                               "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);        \
    } while (0)

#define BA_TRY(expr)                   \
    do {                               \
        int s_ = (expr);               \
        if (s_ != BA_OK) return s_;    \
    } while (0)

// Owning device allocation (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(e == hipErrorOutOfMemory ? BA_ERR_OUT_OF_MEMORY : BA_ERR_HIP, "hipMalloc(%zu bytes): %s",
                        count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return BA_OK;
    }
    int upload(const T* host, size_t count, hipStream_t s) {
        BA_TRY(alloc(count));
        if (count) BA_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        return BA_OK;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Validates a problem view: sizes, camera_block_size 9, index ranges and finite inputs.
int validate_view(const ba_problem_view* v, bool need_params);

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Sets the calling thread's device; returns BA_ERR_NO_DEVICE if it does not exist.  The library's
// first use of a device also creates that device's lane streams (below).
int use_device(int device);

// Streams for the concurrently running lanes of a window batch.  The runtime maps streams onto a
// few hardware queues (GPU_MAX_HW_QUEUES, 4 by default) in the order they are created, and two
// lanes on one queue run one after the other: with one single-problem solver created and destroyed
// before a 512-window batch, its four lanes' own new streams ran it in 24 ms against 20 ms.  So
// kLaneStreams streams are created on each device's first use, before any other stream of the
// library, and kept until the process ends; a batch takes them from this pool and gives them back
// when it is destroyed (a batch finding the pool empty, because another holds its streams, creates
// its own and destroys them).
constexpr int kLaneStreams = 4;
int acquire_lane_streams(int device, int count, std::vector<hipStream_t>* out, std::vector<bool>* pooled);
void release_lane_streams(int device, const std::vector<hipStream_t>& streams, const std::vector<bool>& pooled);

// Allows `kernel` at least `bytes` of dynamic LDS on the current device.  The attribute is process
// wide: it is set under a lock and only when a larger value is asked for, so solvers and servers
// created or stepped on several threads do not call into the runtime for it concurrently.
int allow_dynamic_lds(const void* kernel, int bytes);

// Waits until the work enqueued on `s` so far has finished.  timeout_s <= 0: hipStreamSynchronize.
// Otherwise `ev` is recorded on the stream and polled; after timeout_s seconds without completion
// the wait returns BA_ERR_TIMEOUT naming `what`, instead of blocking the caller for good.
int wait_stream(hipStream_t s, hipEvent_t ev, double timeout_s, const char* what);

}  // namespace mba
