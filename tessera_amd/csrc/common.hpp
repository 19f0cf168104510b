// Shared helpers of the tessera_qa library: HIP error handling, complex
// arithmetic on double2, wavefront reductions, per-thread error text.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tq {

// ---- errors -------------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

#define TQ_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::tq::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")");           \
    } while (0)

#define TQ_LAUNCH_CHECK() TQ_HIP(hipGetLastError())

// ---- device memory ------------------------------------------------------
template <class T>
struct DevBuf {
    T* p = nullptr;
    std::size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(std::size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(std::size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(-2, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Device buffers of one call. Without a region each take() allocates its own
// buffer, freed with the arena; with a region (base, cap) every take() carves
// 256-byte aligned space from it and nothing is allocated (a request beyond
// the region throws TESSERA_ERR_TOO_LARGE).
struct DevArena {
    unsigned char* base = nullptr;
    std::size_t cap = 0, used = 0;
    std::vector<DevBuf<unsigned char>> owned;
    DevArena() = default;
    DevArena(void* region, std::size_t bytes) : base(static_cast<unsigned char*>(region)), cap(bytes) {}
    template <class T>
    T* take(std::size_t count) {
        const std::size_t bytes = count * sizeof(T);
        if (!base) {
            owned.emplace_back(bytes);
            return reinterpret_cast<T*>(owned.back().p);
        }
        const std::size_t at = (used + 255) & ~(std::size_t)255;
        if (at > cap || bytes > cap - at)
            throw Error(-2, "static scratch of " + std::to_string(cap) + " bytes exceeded (" +
                                std::to_string(at + bytes) + " needed)");
        used = at + bytes;
        return reinterpret_cast<T*>(base + at);
    }
};

// ---- complex on double2 ---------------------------------------------------
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double cnorm(double2 a) { return a.x * a.x + a.y * a.y; }

// 4x4 two-site gate, row-major G[r][c] acting on |s_i s_{i+1}> = 2*s_i + s_{i+1}.
struct Gate4 {
    double2 g[16];
};

// ---- wavefront reductions (wave64) --------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace tq
