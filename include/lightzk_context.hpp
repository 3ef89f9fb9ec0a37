// Reference-shaped C++ API over libhzk.so: shared context and error handling.
//
// The reference's operators are CUDA templates launched on the legacy default
// stream with no error reporting (include/operator/*.cuh). The shims in
// include/operator/ and include/protocol/ keep their names and parameter
// lists but call the C ABI of include/hzk.h: they run on one context per
// thread (device chosen with lightzk::set_device before first use, default
// 0), wait for all earlier work on the device before they start (as the
// default stream would) and for their own work before they return, and throw
// lightzk::error with the library's message instead of failing silently.
#pragma once
#include <stdexcept>
#include <string>
#include "hzk.h"

namespace lightzk {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline int& device_slot() {
    static thread_local int dev = 0;
    return dev;
}
inline void set_device(int device) { device_slot() = device; }

struct ctx_holder {
    hzk_ctx* ctx = nullptr;
    ~ctx_holder() { hzk_ctx_destroy(ctx); }
};

inline hzk_ctx* context() {
    static thread_local ctx_holder h;
    if (!h.ctx) {
        int rc = hzk_ctx_create(device_slot(), nullptr, &h.ctx);
        if (rc) throw error(rc, "hzk_ctx_create failed (no usable GPU?)");
    }
    return h.ctx;
}

inline void check(int rc) {
    if (rc) throw error(rc, hzk_last_error(context()));
}

// run f(ctx) ordered after all earlier device work, and wait for it
template <class Fn>
inline void run(Fn f) {
    hzk_ctx* c = context();
    check(hzk_device_sync(c));
    check(f(c));
    check(hzk_sync(c));
}

}  // namespace lightzk
