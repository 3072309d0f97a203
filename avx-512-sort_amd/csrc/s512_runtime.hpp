// s512_runtime.hpp -- host-side runtime shared by the entry points: error
// state, per-device grow-only workspaces, pinned staging, launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include <vector>
#include <cstring>
#include <cstdio>
#include "../../include/s512.h"

namespace s512 {

void set_error(const char* fmt, ...);
const char* last_error();

#define S512_HIP(expr)                                                                      \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::s512::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                              __LINE__);                                                    \
            return S512_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

#define S512_CHECK_LAUNCH() S512_HIP(hipGetLastError())

// A grow-only device buffer.  Growing synchronises the stream first so that
// kernels still reading the old allocation have finished.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    s512_status reserve(size_t bytes, hipStream_t st) {
        if (bytes <= cap) return S512_OK;
        if (ptr) {
            S512_HIP(hipStreamSynchronize(st));
            S512_HIP(hipFree(ptr));
            ptr = nullptr;
            cap = 0;
        }
        size_t want = bytes + bytes / 8 + 4096;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return S512_ERR_NOMEM;
        }
        cap = want;
        return S512_OK;
    }
    void release() {  // caller has synchronised the device
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct HostBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    s512_status reserve(size_t bytes, hipStream_t st) {
        if (bytes <= cap) return S512_OK;
        if (ptr) {
            S512_HIP(hipStreamSynchronize(st));
            S512_HIP(hipHostFree(ptr));
            ptr = nullptr;
            cap = 0;
        }
        size_t want = bytes + bytes / 8 + 4096;
        hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error("hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return S512_ERR_NOMEM;
        }
        cap = want;
        return S512_OK;
    }
    void release() {  // caller has synchronised the device
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

// Everything one device needs between calls.  Calls on one device are
// serialised by `mu` (the workspace is shared).
struct DeviceState {
    std::mutex mu;
    DevBuf data;      // second key array (and value array for key/value sorts)
    DevBuf arena;     // per-level metadata, histograms, lists
    DevBuf stage;     // caller-host-pointer staging
    DevBuf learn;     // learned partitioner: sample, training state, table
    DevBuf ids;       // bucket of every key of a partition level (u16)
    HostBuf pinned;   // host<->device metadata staging
    HostBuf pinned2;
    HostBuf lpinned;  // learned: the training loss, written by the training stream
    hipStream_t lstream = nullptr;  // learned: training runs here, beside the first level
    hipEvent_t lready = nullptr, ltrained = nullptr;
    std::vector<hipEvent_t> events;  // profiling event pool
};

// Per-call kernel timing with HIP events recorded on the call's stream.
struct Profiler {
    bool on = false;
    hipStream_t st = nullptr;
    DeviceState* ds = nullptr;
    size_t used = 0;
    std::vector<int> phase;
    hipEvent_t ev(size_t i) {
        while (ds->events.size() <= i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ds->events.push_back(e);
        }
        return ds->events[i];
    }
    void begin(int ph) {
        if (!on) return;
        hipEvent_t e = ev(used);
        if (e) (void)hipEventRecord(e, st);
        used++;
        phase.push_back(ph);
    }
    void end() {
        if (!on) return;
        hipEvent_t e = ev(used);
        if (e) (void)hipEventRecord(e, st);
        used++;
    }
    // after the stream has been synchronised
    void collect(double* ms, uint32_t* launches) {
        if (!on) return;
        for (size_t k = 0; k < phase.size(); ++k) {
            float t = 0.f;
            hipEvent_t a = ev(2 * k), b = ev(2 * k + 1);
            if (a && b && hipEventElapsedTime(&t, a, b) == hipSuccess) ms[phase[k]] += t;
            launches[phase[k]] += 1;
        }
    }
};
DeviceState& device_state(int dev);
// Frees the cached buffers of the calling device (each module its own).
void release_device_state(int dev);
void release_host_stage(int dev);
void release_select_scratch(int dev);
s512_status current_device(int* dev);

// Bump allocator over one DevBuf for a level's metadata.
struct Arena {
    char* base = nullptr;
    size_t off = 0;
    template <class X>
    X* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        X* p = reinterpret_cast<X*>(base + off);
        off += count * sizeof(X);
        return p;
    }
};

struct SizeCounter {
    size_t off = 0;
    template <class X>
    void take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        off += count * sizeof(X);
    }
};

// Is p device-accessible memory (device or managed allocation)?
bool is_device_pointer(const void* p);

}  // namespace s512
