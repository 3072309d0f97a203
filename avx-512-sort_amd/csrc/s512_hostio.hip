// s512_hostio.hip -- pipelined staging of host arrays (see s512_hostio.hpp).
#include "s512_hostio.hpp"
#include <algorithm>
#include <condition_variable>
#include <thread>

namespace s512 {
namespace {

constexpr size_t CHUNK = 32u << 20;  // bytes per pinned ring buffer
constexpr int NRING = 3;             // pinned ring buffers
constexpr size_t PART = 2u << 20;    // bytes per worker copy task
constexpr size_t SMALL = 4u << 20;   // below this a plain hipMemcpy is used
constexpr int MAX_DEV = 64;

// Persistent worker threads for host memcpy of one chunk at a time; the
// calling thread takes part too.  Every job field is read and written under
// the mutex, only the memcpy itself runs outside it.  Created once, never
// destroyed (its threads are detached and wait on the condition variable).
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* p = new CopyPool();
        return *p;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        if (bytes <= PART || nthreads_ == 0) {
            memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(job_mu_);  // one job at a time
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char*)dst;
            src_ = (const char*)src;
            bytes_ = bytes;
            parts_ = (bytes + PART - 1) / PART;
            next_ = 0;
            finished_ = 0;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return finished_ == parts_; });
    }

private:
    CopyPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        nthreads_ = (int)std::min<unsigned>(hw > 1 ? hw - 1 : 0, 8u);
        for (int i = 0; i < nthreads_; ++i) std::thread([this] { loop(); }).detach();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            work();
        }
    }
    void work() {
        for (;;) {
            char* d;
            const char* s;
            size_t len;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (next_ >= parts_) return;
                const size_t off = next_ * PART;
                d = dst_ + off;
                s = src_ + off;
                len = std::min(PART, bytes_ - off);
                ++next_;
            }
            memcpy(d, s, len);
            std::lock_guard<std::mutex> lk(mu_);
            if (++finished_ == parts_) done_cv_.notify_all();
        }
    }
    int nthreads_ = 0;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, parts_ = 0, next_ = 0, finished_ = 0;
    uint64_t gen_ = 0;
};

}  // namespace

struct HostStage::State {
    std::mutex mu;
    bool ready = false;
    DevBuf dev[3];
    HostBuf ring;
    hipStream_t cs = nullptr;
    hipEvent_t ev[NRING] = {};

    // copy stream, events and pinned ring: created by the first large copy
    s512_status init() {
        if (ready) return S512_OK;
        // a stream of default priority that does not synchronise with the null stream
        S512_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        for (int k = 0; k < NRING; ++k) S512_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        if (s512_status rc = ring.reserve(NRING * CHUNK, cs)) return rc;
        ready = true;
        return S512_OK;
    }
    char* buf(int k) { return (char*)ring.ptr + (size_t)k * CHUNK; }

    // host -> device: worker threads fill ring buffer k while the DMA of the
    // previous chunk runs; a buffer is refilled only after its DMA completed
    s512_status h2d(char* d, const char* h, size_t bytes) {
        if (s512_status rc = init()) return rc;
        const size_t nch = (bytes + CHUNK - 1) / CHUNK;
        for (size_t i = 0; i < nch; ++i) {
            const int k = (int)(i % NRING);
            if (i >= NRING) S512_HIP(hipEventSynchronize(ev[k]));
            const size_t len = std::min(CHUNK, bytes - i * CHUNK);
            CopyPool::get().copy(buf(k), h + i * CHUNK, len);
            S512_HIP(hipMemcpyAsync(d + i * CHUNK, buf(k), len, hipMemcpyHostToDevice, cs));
            S512_HIP(hipEventRecord(ev[k], cs));
        }
        S512_HIP(hipStreamSynchronize(cs));
        return S512_OK;
    }
    // device -> host: the DMA of the next NRING chunks is in flight while
    // worker threads empty the oldest ring buffer into the caller's memory
    s512_status d2h(char* h, const char* d, size_t bytes) {
        if (s512_status rc = init()) return rc;
        const size_t nch = (bytes + CHUNK - 1) / CHUNK;
        auto issue = [&](size_t j) -> s512_status {
            const int k = (int)(j % NRING);
            const size_t len = std::min(CHUNK, bytes - j * CHUNK);
            S512_HIP(hipMemcpyAsync(buf(k), d + j * CHUNK, len, hipMemcpyDeviceToHost, cs));
            S512_HIP(hipEventRecord(ev[k], cs));
            return S512_OK;
        };
        for (size_t j = 0; j < std::min<size_t>(NRING, nch); ++j)
            if (s512_status rc = issue(j)) return rc;
        for (size_t i = 0; i < nch; ++i) {
            const int k = (int)(i % NRING);
            S512_HIP(hipEventSynchronize(ev[k]));
            const size_t len = std::min(CHUNK, bytes - i * CHUNK);
            CopyPool::get().copy(h + i * CHUNK, buf(k), len);
            if (i + NRING < nch)
                if (s512_status rc = issue(i + NRING)) return rc;
        }
        return S512_OK;
    }
};

static std::mutex g_states_mu;
static HostStage::State* g_states[MAX_DEV];

HostStage::HostStage() {
    int dev = 0;
    if ((rc_ = current_device(&dev))) return;
    if (dev < 0 || dev >= MAX_DEV) {
        set_error("device %d out of range", dev);
        rc_ = S512_ERR_INVALID;
        return;
    }
    {
        std::lock_guard<std::mutex> lk(g_states_mu);
        if (!g_states[dev]) g_states[dev] = new State();
        s_ = g_states[dev];
    }
    s_->mu.lock();
}

HostStage::~HostStage() {
    if (!s_) return;
    // a failed call may leave work on its staged buffers: finish it before
    // another call can grow (free) them
    (void)hipStreamSynchronize(nullptr);
    s_->mu.unlock();
}

s512_status HostStage::in(int slot, const void* host, size_t bytes, void** dev, bool copy_in) {
    if (rc_) return rc_;
    if (slot < 0 || slot > 2) {
        set_error("staging slot %d", slot);
        return S512_ERR_INVALID;
    }
    if (s512_status rc = s_->dev[slot].reserve(std::max<size_t>(bytes, 1), s_->cs)) return rc;
    *dev = s_->dev[slot].ptr;
    if (!copy_in || bytes == 0) return S512_OK;
    if (bytes <= SMALL) {
        S512_HIP(hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice));
        return S512_OK;
    }
    return s_->h2d((char*)*dev, (const char*)host, bytes);
}

s512_status HostStage::out(int slot, void* host, size_t bytes) {
    if (rc_) return rc_;
    // the device work on the slot ran on the null stream
    S512_HIP(hipStreamSynchronize(nullptr));
    if (bytes == 0) return S512_OK;
    if (bytes <= SMALL) {
        S512_HIP(hipMemcpy(host, s_->dev[slot].ptr, bytes, hipMemcpyDeviceToHost));
        return S512_OK;
    }
    return s_->d2h((char*)host, (const char*)s_->dev[slot].ptr, bytes);
}

}  // namespace s512

namespace s512 {
void release_host_stage(int dev) {
    if (dev < 0 || dev >= MAX_DEV) return;
    HostStage::State* st;
    {
        std::lock_guard<std::mutex> lk(g_states_mu);
        st = g_states[dev];
    }
    if (!st) return;
    std::lock_guard<std::mutex> lk(st->mu);
    (void)hipDeviceSynchronize();
    for (auto& b : st->dev) b.release();
    // the pinned ring, stream and events stay: they are created once
}
}  // namespace s512
