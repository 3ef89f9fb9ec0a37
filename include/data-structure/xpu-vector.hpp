// Mirrored host/device buffer with the interface of the reference's
// xpu::vector (include/data-structure/xpu-vector.cuh:5-101): the same policy
// names, constructor, allocate(), p(), operator[], store(), load() overloads
// and clear(). Device memory comes from the library (hzk_device_alloc on the
// calling thread's context, include/lightzk_context.hpp); failures throw
// lightzk::error instead of being ignored.
//
// Differences from the reference, each a defect there: copying is disabled
// (the reference's implicit copy frees the same memory twice), moving is
// allowed; allocate() on an allocated vector releases the old storage first
// (the reference leaks it); a default-constructed vector owns nothing and
// destroys cleanly (the reference reads an uninitialised policy).
#pragma once
#include <cassert>
#include <cstddef>
#include <memory>
#include <utility>
#include "../lightzk_context.hpp"

namespace xpu {

enum class mem_policy { host_only, device_only, cross_platform };

template <typename T>
class vector {
    std::unique_ptr<T[]> cpu_;
    void* gpu_ = nullptr;
    size_t size_ = 0;
    mem_policy mp_ = mem_policy::host_only;

    void release() {
        if (gpu_) hzk_device_free(lightzk::context(), gpu_);
        gpu_ = nullptr;
        cpu_.reset();
        size_ = 0;
    }

public:
    T& operator[](size_t idx) { return cpu_[idx]; }
    const T& operator[](size_t idx) const { return cpu_[idx]; }
    void* p() {
        assert(mp_ != mem_policy::host_only);
        return gpu_;
    }
    size_t size() const { return size_; }
    mem_policy policy() const { return mp_; }

    vector() = default;
    vector(size_t size, mem_policy mp = mem_policy::host_only) { allocate(size, mp); }
    vector(const vector&) = delete;
    vector& operator=(const vector&) = delete;
    vector(vector&& o) noexcept : cpu_(std::move(o.cpu_)), gpu_(o.gpu_), size_(o.size_), mp_(o.mp_) {
        o.gpu_ = nullptr;
        o.size_ = 0;
    }
    vector& operator=(vector&& o) noexcept {
        if (this != &o) {
            release();
            cpu_ = std::move(o.cpu_);
            gpu_ = o.gpu_, size_ = o.size_, mp_ = o.mp_;
            o.gpu_ = nullptr, o.size_ = 0;
        }
        return *this;
    }
    ~vector() { release(); }

    void allocate(size_t size, mem_policy mp = mem_policy::host_only) {
        release();
        size_ = size, mp_ = mp;
        if (mp != mem_policy::device_only) cpu_.reset(new T[size ? size : 1]);
        if (mp != mem_policy::host_only)
            lightzk::check(hzk_device_alloc(lightzk::context(), (size ? size : 1) * sizeof(T), &gpu_));
    }

    // host -> device, whole vector
    void store() {
        assert(mp_ == mem_policy::cross_platform);
        lightzk::check(hzk_device_sync(lightzk::context()));
        lightzk::check(hzk_memcpy_h2d(lightzk::context(), gpu_, cpu_.get(), size_ * sizeof(T)));
    }
    // device -> host, whole vector or its first `length` elements
    void load() { load(size_); }
    void load(size_t length) {
        assert(mp_ == mem_policy::cross_platform && length <= size_);
        lightzk::check(hzk_device_sync(lightzk::context()));
        lightzk::check(hzk_memcpy_d2h(lightzk::context(), cpu_.get(), gpu_, length * sizeof(T)));
    }
    // `length` elements from any device address into the host copy
    void load(void* src, size_t length) {
        assert(mp_ != mem_policy::device_only && length <= size_);
        lightzk::check(hzk_device_sync(lightzk::context()));
        lightzk::check(hzk_memcpy_d2h(lightzk::context(), cpu_.get(), src, length * sizeof(T)));
    }
    // zero the device copy and wait for it
    void clear() {
        assert(mp_ != mem_policy::host_only);
        lightzk::run([&](hzk_ctx* c) { return hzk_memset(c, gpu_, 0, size_ * sizeof(T)); });
    }
};

}  // namespace xpu
