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

// DPP move of a double within 16-lane rows (CTRL is a DPP control code).
template <int CTRL>
__device__ inline double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes, result in every lane: quad_perm and row_ror inside each
// 16-lane row (DPP, no LDS round trip), then two xor shuffles across rows.
__device__ inline double wave_sum_dpp(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x124>(v);  // row_ror 4
    v += dpp_d<0x128>(v);  // row_ror 8
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

}  // namespace tq
