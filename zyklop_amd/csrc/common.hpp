// Host-side plumbing shared by the library: HIP error handling (errors become
// C++ exceptions inside the library and status codes at the C-ABI), device
// buffers, and small launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zk_abi.h"

#include <stdexcept>
#include <string>

namespace zk {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes (ZK_OK, ZK_E_ARG, ...) are the C-ABI's, from zk_abi.h.

#define ZK_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t _e = (x);                                                                      \
        if (_e != hipSuccess)                                                                     \
            throw ::zk::Error(ZK_E_HIP, std::string(#x) + ": " + hipGetErrorString(_e) +    \
                                                  " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define ZK_REQUIRE(cond, code, msg)                         \
    do {                                                    \
        if (!(cond)) throw ::zk::Error((code), (msg));      \
    } while (0)

// Device allocation: owning (hipMalloc), or borrowed from memory the caller
// manages (borrow()), which never grows and is never freed here.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), borrowed(o.borrowed) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes; borrowed = o.borrowed;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void borrow(void* q, size_t n) {
        release();
        p = q;
        bytes = n;
        borrowed = true;
    }
    void alloc(size_t n) {
        if (borrowed) throw Error(ZK_E_INTERNAL, "a borrowed device buffer cannot grow");
        release();
        if (n == 0) n = 256;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(ZK_E_NOMEM, "hipMalloc(" + std::to_string(n) + " bytes) failed: " + hipGetErrorString(e));
        }
        bytes = n;
    }
    // grow-only reallocation
    void ensure(size_t n) {
        if (n > bytes) alloc(n);
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Device bytes a sizing pass (DevBuf::ensure on every buffer) would add: each
// requirement is counted against the buffer's current size, or in full when
// `fresh` (a workspace not yet created). Admission control uses it to know
// whether queueing work would allocate.
struct GrowCount {
    bool fresh = false;
    size_t bytes = 0;
    void need(const DevBuf& d, size_t n) {
        const size_t cur = fresh ? 0 : d.bytes;
        if (n > cur) bytes += n - cur;
    }
};

// Owning pinned host allocation (hipHostMalloc).
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void ensure(size_t n) {
        if (n <= bytes) return;
        release();
        ZK_HIP(hipHostMalloc(&p, n, hipHostMallocDefault));
        bytes = n;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

inline int ilog2(uint64_t x) {
    int r = -1;
    while (x) { x >>= 1; r++; }
    return r;
}

}  // namespace zk
