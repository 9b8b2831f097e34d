// Support of the C++ interface headers (graph.hpp, coarsen.hpp, ...): status
// codes of the C ABI turned back into the reference's exceptions, and the
// per-thread session context that plays the role of the reference's
// thread-local ScratchPoolSession (include/detail/scratch_pool.hpp:115-130):
// every device object made on a thread lives in that thread's context,
// which stays alive while any object made in it does.
#pragma once

#include "gpart_mi355x.h"

#include <hip/hip_runtime_api.h>

#include <memory>
#include <new>
#include <stdexcept>
#include <string>

namespace gpart {
namespace detail {

// GP_ERR_INVALID -> std::invalid_argument, GP_ERR_RUNTIME -> std::runtime_error
// (overflow_error where the caller says so), GP_ERR_HIP -> std::runtime_error
// (CUDA_CHECK, include/check.hpp:8-18), GP_ERR_MEMORY -> std::bad_alloc.
inline void check(int status, bool runtime_is_overflow = false) {
    if (status == GP_OK) return;
    const std::string what = gp_last_error();
    switch (status) {
    case GP_ERR_INVALID:
        throw std::invalid_argument(what);
    case GP_ERR_MEMORY:
        throw std::bad_alloc();
    case GP_ERR_RUNTIME:
        if (runtime_is_overflow) throw std::overflow_error(what);
        throw std::runtime_error(what);
    default:
        throw std::runtime_error(what);
    }
}

class Session {
public:
    explicit Session(int device) : device_(device) { check(gp_context_create(device, &ctx_)); }
    ~Session() { gp_context_destroy(ctx_); }
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;
    gp_context* get() const { return ctx_; }
    int device() const { return device_; }

private:
    gp_context* ctx_ = nullptr;
    int device_ = 0;
};

// The calling thread's context on the current HIP device, created on first use.
inline std::shared_ptr<Session> session() {
    thread_local std::shared_ptr<Session> current;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) device = 0;
    if (!current || current->device() != device) current = std::make_shared<Session>(device);
    return current;
}

}  // namespace detail
}  // namespace gpart
