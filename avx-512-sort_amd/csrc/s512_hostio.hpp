// s512_hostio.hpp -- host-pointer entry points: caller arrays in ordinary
// (pageable) host memory staged through the device.
//
// The reference sorts arrays in host memory (sort512.hpp:5960), so every
// drop-in entry point accepts a host pointer.  A plain pageable hipMemcpy
// moves 30 GB/s host-to-device and under 18 GB/s back on MI355X, against
// 57 GB/s for pinned memory.  Here a host array moves in chunks through a
// ring of pinned buffers: worker threads copy chunk i+1 between the caller's
// memory and a pinned buffer while the DMA engine moves chunk i, on a copy
// stream of its own.  The device copy lives in a grow-only per-device buffer
// (one per staged array slot), kept between calls.
#pragma once
#include "s512_runtime.hpp"

namespace s512 {

// Holds the calling device's host-staging state (buffers, copy stream) for
// the guard's lifetime; staged calls on one device are serialised.
class HostStage {
public:
    HostStage();
    ~HostStage();
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    s512_status status() const { return rc_; }
    // Device copy of `bytes` bytes of host memory in slot 0..2 (keys, values,
    // offsets); `copy_in` false only reserves the device buffer.
    s512_status in(int slot, const void* host, size_t bytes, void** dev, bool copy_in = true);
    // Device slot -> host memory.
    s512_status out(int slot, void* host, size_t bytes);

    struct State;  // per-device buffers and copy stream (s512_hostio.hip)

private:
    State* s_ = nullptr;
    s512_status rc_ = S512_OK;
};

// Run fn(device pointer, stream) on the n elements at a: in place for device
// memory, on a staged device copy (copied back when fn succeeds) for host
// memory.
template <class T, class Fn>
s512_status with_staged_array(T* a, size_t n, Fn fn) {
    if (!a) {
        set_error("null array pointer");
        return S512_ERR_INVALID;
    }
    if (is_device_pointer(a)) return fn(a, (hipStream_t)0);
    HostStage hs;
    void* d = nullptr;
    if (s512_status rc = hs.in(0, a, n * sizeof(T), &d)) return rc;
    if (s512_status rc = fn((T*)d, (hipStream_t)0)) return rc;
    return hs.out(0, a, n * sizeof(T));
}

// The same for a key array and an int32 value array of n elements each
// (slots 0 and 1, under one guard: a guard must never be nested on one device).
template <class K, class Fn>
s512_status with_staged_pair(K* k, int32_t* v, size_t n, Fn fn) {
    if (!k || !v) {
        set_error("null array pointer");
        return S512_ERR_INVALID;
    }
    const bool kdev = is_device_pointer(k), vdev = is_device_pointer(v);
    if (kdev && vdev) return fn(k, v, (hipStream_t)0);
    HostStage hs;
    void* dk = k;
    void* dv = v;
    const size_t kbytes = n * sizeof(K), vbytes = n * sizeof(int32_t);
    if (!kdev)
        if (s512_status rc = hs.in(0, k, kbytes, &dk)) return rc;
    if (!vdev)
        if (s512_status rc = hs.in(1, v, vbytes, &dv)) return rc;
    if (s512_status rc = fn((K*)dk, (int32_t*)dv, (hipStream_t)0)) return rc;
    if (!kdev)
        if (s512_status rc = hs.out(0, k, kbytes)) return rc;
    if (!vdev)
        if (s512_status rc = hs.out(1, v, vbytes)) return rc;
    return S512_OK;
}

}  // namespace s512
