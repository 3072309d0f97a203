// s512_runtime.hip -- error state, device workspaces, pointer classification.
#include "s512_runtime.hpp"
#include <cstdarg>
#include <map>
#include <memory>

namespace s512 {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }

DeviceState& device_state(int dev) {
    static std::mutex map_mu;
    static std::map<int, std::unique_ptr<DeviceState>> states;
    std::lock_guard<std::mutex> lk(map_mu);
    auto& p = states[dev];
    if (!p) p.reset(new DeviceState());
    return *p;
}

void release_device_state(int dev) {
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    (void)hipDeviceSynchronize();
    ds.data.release();
    ds.arena.release();
    ds.stage.release();
    ds.learn.release();
    ds.ids.release();
    ds.pinned.release();
    ds.pinned2.release();
}

s512_status current_device(int* dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "count is 0" : hipGetErrorString(e));
        return S512_ERR_NODEVICE;
    }
    S512_HIP(hipGetDevice(dev));
    return S512_OK;
}

bool is_device_pointer(const void* p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // plain host memory: clear the sticky error
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

}  // namespace s512

extern "C" const char* s512_last_error(void) { return s512::last_error(); }
extern "C" const char* s512_version(void) { return "s512-mi355x 0.1 (gfx950)"; }

extern "C" s512_status s512_release(void) {
    int dev;
    if (s512_status rc = s512::current_device(&dev)) return rc;
    s512::release_host_stage(dev);
    s512::release_select_scratch(dev);
    s512::release_device_state(dev);
    return S512_OK;
}
