// Device plumbing of the pomerol-compatible API: the process-wide gfx950 context and owning
// device arrays.  Every GPU call goes through the C ABI in pomx_c.h; a failure raises
// std::runtime_error carrying pomx_last_error().  There is no CPU fallback: without a usable
// gfx950 device, constructing the context throws.
#pragma once

#include "../pomx_c.h"
#include "Misc.hpp"

#include <memory>
#include <vector>

namespace Pomerol {
namespace detail {

/// The context of the current host thread (created on first use).
pomx_ctx* gpu();
/// Throws std::runtime_error with the context's message if rc != POMX_OK.
void check(int rc, char const* what);

/// Owning device array.
template <typename T> class DeviceArray {
    T* p_ = nullptr;
    std::size_t n_ = 0;

public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) : n_(n) {
        void* p = nullptr;
        check(pomx_malloc(gpu(), sizeof(T) * (n ? n : 1), &p), "device allocation");
        p_ = static_cast<T*>(p);
    }
    DeviceArray(DeviceArray const&) = delete;
    DeviceArray& operator=(DeviceArray const&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if(this != &o) {
            reset();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DeviceArray() { reset(); }
    void reset() {
        if(p_)
            pomx_free(gpu(), p_);
        p_ = nullptr;
        n_ = 0;
    }
    /// Take ownership of memory allocated by a pomx_* call.
    static DeviceArray adopt(T* p, std::size_t n) {
        DeviceArray a;
        a.p_ = p;
        a.n_ = n;
        return a;
    }
    T* data() { return p_; }
    T const* data() const { return p_; }
    std::size_t size() const { return n_; }
    void upload(T const* h, std::size_t n, std::size_t offset = 0) {
        check(pomx_h2d(gpu(), p_ + offset, h, sizeof(T) * n), "host to device copy");
    }
    void upload(std::vector<T> const& h) { upload(h.data(), h.size()); }
    void download(T* h, std::size_t n, std::size_t offset = 0) const {
        check(pomx_d2h(gpu(), h, p_ + offset, sizeof(T) * n), "device to host copy");
    }
    std::vector<T> download() const {
        std::vector<T> h(n_);
        download(h.data(), n_);
        return h;
    }
    void zero() { check(pomx_fill_zero(gpu(), p_, sizeof(T) * n_), "device fill"); }
};

template <typename T> using SharedDeviceArray = std::shared_ptr<DeviceArray<T>>;

} // namespace detail
} // namespace Pomerol
