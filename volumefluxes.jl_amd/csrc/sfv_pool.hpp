// sfv_pool.hpp — device memory, streams and pinned step-control blocks kept
// for reuse by handles that a long-running owner creates and destroys (the
// server: one handle per request).  hipFree waits for the whole device, so
// returning a request's buffers to the pool instead keeps the batched steps
// of the other requests running; hipHostMalloc / hipStreamCreate are slow
// enough per request to matter too.
//
// Device blocks are handed out by size class (the byte count rounded up to
// 1/8 of its power of two above 64 KiB, to 4 KiB below), exact class only,
// so a handle holds exactly the class-rounded bytes its owner budgeted.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <mutex>
#include <vector>

namespace sfv {

class DevicePool {
public:
    DevicePool() = default;
    DevicePool(const DevicePool&) = delete;
    DevicePool& operator=(const DevicePool&) = delete;
    ~DevicePool() {
        for (auto& kv : free_) (void)hipFree(kv.second);
        for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
        for (void* p : pinned_) (void)hipHostFree(p);
    }

    static size_t size_class(size_t bytes) {
        if (bytes <= (size_t)64 << 10) return (bytes + 4095) & ~(size_t)4095;
        int e = 0;
        while (((size_t)1 << (e + 1)) <= bytes) ++e;          // 2^e <= bytes < 2^(e+1)
        const size_t step = (size_t)1 << (e - 3);
        return (bytes + step - 1) / step * step;
    }

    // A block of exactly `cls` bytes (a size_class value): a cached one, else
    // a new allocation.  nullptr when the device is out of memory.
    void* get(size_t cls) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.find(cls);
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                cached_ -= cls;
                return p;
            }
        }
        void* p = nullptr;
        return hipMalloc(&p, cls) == hipSuccess ? p : nullptr;
    }
    void put(void* p, size_t cls) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(cls, p);
        cached_ += cls;
    }
    // Free cached blocks, largest first, until at most `keep` bytes stay cached.
    void trim(size_t keep) {
        std::lock_guard<std::mutex> lk(mu_);
        while (cached_ > keep && !free_.empty()) {
            auto it = std::prev(free_.end());
            (void)hipFree(it->second);
            cached_ -= it->first;
            free_.erase(it);
        }
    }
    size_t cached() {
        std::lock_guard<std::mutex> lk(mu_);
        return cached_;
    }

    hipStream_t get_stream() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!streams_.empty()) {
                hipStream_t s = streams_.back();
                streams_.pop_back();
                return s;
            }
        }
        hipStream_t s = nullptr;
        return hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess ? s : nullptr;
    }
    void put_stream(hipStream_t s) {
        std::lock_guard<std::mutex> lk(mu_);
        streams_.push_back(s);
    }

    // Pinned host blocks of one fixed size (the step-control mirror).
    void* get_pinned(size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!pinned_.empty() && bytes <= pinned_bytes_) {
                void* p = pinned_.back();
                pinned_.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        return hipHostMalloc(&p, bytes < pinned_bytes_ ? pinned_bytes_ : bytes, hipHostMallocDefault) == hipSuccess
                   ? p
                   : nullptr;
    }
    void put_pinned(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        pinned_.push_back(p);
    }

private:
    static constexpr size_t pinned_bytes_ = 4096;
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t cached_ = 0;
    std::vector<hipStream_t> streams_;
    std::vector<void*> pinned_;
};

}  // namespace sfv
