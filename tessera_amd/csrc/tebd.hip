// MPS-TEBD engine on gfx950 (upstream cuda/src/mps.cpp, cuda/src/linalg.cpp).
//
// HBM layout: site i is a (Dl, 2, Dr) complex128 tensor stored with fixed
// capacity strides, element (a, s, b) at off[i] + (a*2 + s)*cap[i+1] + b, where
// cap[b] = min(chi, 2^b, 2^(n-b)) bounds bond b. Current bond dimensions live on
// the device (dims[0..n]); every kernel reads them there and the SVD kernel
// writes the kept rank, so a whole Trotter step is enqueued without a single
// host synchronisation.
//
// Two-site update (mps.cpp:41-100), three kernels:
//   k_theta  theta[(a,r1),(r2,b)] = sum G[r1r2][s1s2] sum_c A(a,s1,c) C(c,s2,b),
//            written column-major as M = theta (if 2Dr <= 2Dl) or theta^dagger.
//   k_svd    one-sided Jacobi on the columns of a copy of M (FP64 angles,
//            norms and convergence test), singular values = final column norms,
//            sorted descending; keep = min(chi, k) minus trailing S < 1e-14
//            (mps.cpp:80-82); discarded weight max-accumulated (mps.cpp:83-85).
//   k_recon  A' = U[:, :keep], C' = U_keep^dagger theta  (= diag(S) Vh, mps.cpp:87-99).
// The upstream computes its SVD from the Gram matrix (linalg.cpp:107-172); this
// engine computes the same factorisation accurately, which is what the upstream
// algorithm defines.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include <hip/hip_cooperative_groups.h>

#include "mfma64.hpp"
#include "tebd.hpp"

namespace tq {

namespace {

constexpr int kSvdThreads = 1024;
constexpr int kMaxCols = 512;  // 2 * chi for chi <= 256
// Convergence: a column pair is orthogonal when |x^H y| <= m * eps * |x| |y|,
// the level at which FP64 dot products of length m stop being informative.
__device__ inline double jacobi_tol(int m) { return (double)(m < 8 ? 8 : m) * 2.220446049250313e-16; }

__device__ inline double2 ld_site(const double2* sites, std::size_t off, int stride, int a, int s, int b) {
    return sites[off + ((std::size_t)a * 2 + s) * stride + b];
}

// ---- XCD-local coherence --------------------------------------------------
// The workgroups of one executor group can run on one XCD (workgroups are dealt
// to the 8 XCDs round-robin, blockIdx mod 8; the kernel checks HW_REG_XCC_ID)
// and then share that XCD's L2. Between them a release needs only every wave's
// stores complete in the L2 (s_waitcnt) and an acquire needs only to keep the
// CU's L1 out of the way: data another workgroup wrote since the last
// device-scope acquire is read with device-scope (sc1) loads, which skip the
// L1 and are served by the L2. This avoids the L2 write-back and invalidation
// that every device-scope release and acquire costs (measured on MI355X: 0.8 us
// for such a barrier of 32 workgroups against 6.8 us at device scope).
// gld<XL>: plain load for XL = false, L1-bypassing load for XL = true.
template <bool XL>
__device__ __forceinline__ double gld(const double* p) {
    if constexpr (XL) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool XL>
__device__ __forceinline__ int gld(const int* p) {
    if constexpr (XL) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool XL>
__device__ __forceinline__ double2 gld(const double2* p) {
    if constexpr (XL) {
        const double* q = reinterpret_cast<const double*>(p);
        return make_double2(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                            __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else {
        return *p;
    }
}
__device__ __forceinline__ unsigned xcc_id() {
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xfu;
}

// ---- single-site half layer (mps.cpp:313-322) -----------------------------
// exp(-i dt/2 g), g = [[-hz, -A], [-A, hz]], hz = B h_i (+ A bias_i), closed form.
__global__ void __launch_bounds__(256) k_single_layer(double2* sites, const std::size_t* off, const int* cap,
                                                      const int* dims, const double* h, const double* bias,
                                                      double A, double B, double dt, StepParams sp) {
    const int i = blockIdx.x;
    if (sp.ab) {  // replayed step: this step's schedule values
        const int k = sp.par[0];
        A = sp.ab[2 * k];
        B = sp.ab[2 * k + 1];
    }
    const int Dl = dims[i], Dr = dims[i + 1], stride = cap[i + 1];
    const double hz = B * h[i] + (bias ? A * bias[i] : 0.0);
    const double r = sqrt(A * A + hz * hz);
    double2 u00 = make_double2(1, 0), u01 = make_double2(0, 0), u11 = make_double2(1, 0);
    if (r > 0.0) {
        double sn, cs;
        sincos(0.5 * dt * r, &sn, &cs);
        u00 = make_double2(cs, sn * hz / r);
        u01 = make_double2(0.0, sn * A / r);
        u11 = make_double2(cs, -sn * hz / r);
    }
    double2* t = sites + off[i];
    for (int idx = threadIdx.x; idx < Dl * Dr; idx += blockDim.x) {
        const int a = idx / Dr, b = idx % Dr;
        double2* p0 = t + ((std::size_t)a * 2 + 0) * stride + b;
        double2* p1 = t + ((std::size_t)a * 2 + 1) * stride + b;
        const double2 v0 = *p0, v1 = *p1;
        *p0 = cadd(cmul(u00, v0), cmul(u01, v1));
        *p1 = cadd(cmul(u01, v0), cmul(u11, v1));
    }
}

// ---- theta (mps.cpp:49-70) on the matrix cores ------------------------------
// P = A C with A = site p as a (2 Dl x mid) matrix (rows (a, s1), row stride
// cap[p+1]) and C = site p+1 as a (mid x 2 Dr) matrix (columns (s2, b), row
// stride 2 cap[p+2]): complex FP64 MFMA products, K = mid in steps of 4. A
// wavefront's tile covers a = a0..a0+7 (both s1) and b = b0..b0+15 in two
// 16 x 16 accumulators (s2 = 0, 1); its tile rows are ordered so that every lane
// ends up holding all four P[(a, s1), (s2, b)] of two values of a, and the 4x4
// gate is applied in registers: theta[(a,r1),(r2,b)] = sum G[r1r2][s1s2] P.
// theta is written column-major as M = theta (2 Dr <= 2 Dl) or theta^dagger,
// and copied to W (the SVD's working matrix). Tiles are dealt to waves
// `wave`, `wave + nwaves`, ...
__device__ void theta_mfma(const double2* __restrict__ sites, const std::size_t* off, const int* cap,
                           const int* dims, int p, const Gate4& g, double2* __restrict__ M,
                           double2* __restrict__ W, int ld, long long wave, long long nwaves,
                           const double2* zzc = nullptr) {
    const int Dl = dims[p], mid = dims[p + 1], Dr = dims[p + 2];
    const int sA = cap[p + 1], sC = cap[p + 2];
    const double2* A = sites + off[p];
    const double2* C = sites + off[p + 1];
    const int lane = (int)(threadIdx.x & 63), lg = lane >> 4, lc = lane & 15;
    const int nta = (Dl + 7) / 8, ntb = (Dr + 15) / 16;
    const bool tall = Dr <= Dl;
    const double2 z = make_double2(0, 0);
    // the gate in LDS: the epilogue reads its coefficients where it needs them
    // (kept in registers across the tile loop they would crowd out the products)
    __shared__ double2 gl[16];
    if (threadIdx.x == 0) {
        if (zzc) {  // exp(-i theta Z(x)Z) from the step's phasor (cos theta, sin theta)
            const double2 cs = *zzc;
#pragma unroll
            for (int q = 0; q < 16; ++q) gl[q] = make_double2(0, 0);
            gl[0] = make_double2(cs.x, -cs.y);
            gl[5] = make_double2(cs.x, cs.y);
            gl[10] = make_double2(cs.x, cs.y);
            gl[15] = make_double2(cs.x, -cs.y);
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) gl[q] = g.g[q];
        }
    }
    __syncthreads();
    for (long long t = wave; t < (long long)nta * ntb; t += nwaves) {
        const int a0 = (int)(t / ntb) * 8, b0 = (int)(t % ntb) * 16;
        // operand row lc = tile row r <-> (a, s1) with r = g + 4 i, a = a0 + 2 g + i / 2, s1 = i % 2
        const int ra = a0 + 2 * (lc & 3) + (lc >> 3), rs = (lc >> 2) & 1;
        const int bb = b0 + lc;
        const bool arow = ra < Dl, bcol = bb < Dr;
        const double2* ap = A + (std::size_t)(ra * 2 + rs) * sA;
        CTile t0, t1;
        t0.zero();
        t1.zero();
        for (int c0 = 0; c0 < mid; c0 += 8) {
            double2 av[2], b0v[2], b1v[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {  // two K steps of loads in flight
                const int c = c0 + 4 * u + lg;
                const bool ok = c < mid;
                av[u] = (ok && arow) ? ap[c] : z;
                b0v[u] = (ok && bcol) ? C[(std::size_t)(c * 2 + 0) * sC + bb] : z;
                b1v[u] = (ok && bcol) ? C[(std::size_t)(c * 2 + 1) * sC + bb] : z;
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                t0.mac(av[u], b0v[u]);
                t1.mac(av[u], b1v[u]);
            }
        }
        if (!bcol) continue;
        // elements 2h, 2h+1 hold a = a0 + 2 lg + h with s1 = 0, 1
        auto emit = [&](int a, double2 p00, double2 p01, double2 p10, double2 p11) {  // p<s1><s2>
            if (a >= Dl) return;
#pragma unroll
            for (int r1 = 0; r1 < 2; ++r1)
#pragma unroll
                for (int r2 = 0; r2 < 2; ++r2) {
                    asm volatile("" ::: "memory");  // re-read the coefficients here
                    const double2* gq = gl + (r1 * 2 + r2) * 4;
                    const double2 acc =
                        cadd(cadd(cmul(gq[0], p00), cmul(gq[1], p01)), cadd(cmul(gq[2], p10), cmul(gq[3], p11)));
                    const int row = a * 2 + r1, col = r2 * Dr + bb;
                    const std::size_t at = tall ? (std::size_t)col * ld + row : (std::size_t)row * ld + col;
                    const double2 v = tall ? acc : cconj(acc);
                    M[at] = v;
                    W[at] = v;
                }
        };
        emit(a0 + 2 * lg, t0.get(0), t1.get(0), t0.get(1), t1.get(1));
        emit(a0 + 2 * lg + 1, t0.get(2), t1.get(2), t0.get(3), t1.get(3));
    }
}

// Number of theta_mfma tiles for bond dimensions up to Dl x Dr.
inline long long theta_tiles_for(int Dl, int Dr) { return (long long)((Dl + 7) / 8) * ((Dr + 15) / 16); }

// theta of the per-update path: one tile per wavefront, grid sized for the
// capacities of the bonds (tiles outside the current dimensions do nothing).
__global__ void __launch_bounds__(256) k_theta(const double2* __restrict__ sites, const std::size_t* off,
                                               const int* cap, const int* dims, int p, Gate4 g,
                                               double2* __restrict__ M, double2* __restrict__ W, int ld,
                                               StepParams sp, int edge) {
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    // replayed step with a ZZ gate: the gate of this step from the phasor table
    const double2* zzc = (sp.zz && edge >= 0) ? sp.zz + (std::size_t)sp.par[0] * sp.m + edge : nullptr;
    theta_mfma(sites, off, cap, dims, p, g, M, W, ld, wave, nwaves, zzc);
}

// Advances a replayed step's index (the last node of a captured Trotter step).
__global__ void k_step_advance(int* par) { par[0] += 1; }

// ---- one-sided Jacobi SVD, one workgroup -----------------------------------
__device__ inline void rr_pair(int r, int k, int cp, int& i, int& j) {
    if (k == 0) {
        i = r;
        j = cp - 1;
    } else {
        i = (r + k) % (cp - 1);
        j = (r - k + (cp - 1)) % (cp - 1);
    }
}

// Sort key of a column norm for the stable descending rank: NaN ranks below
// every number, so the ranks always form a permutation of the columns (the
// permutation indexes the working matrix) whatever the data.
__device__ inline double rank_key(double x) { return x == x ? x : -1.0; }

// Rows held per lane for m <= 512.
constexpr int kRowsPerLane = 8;

// Pair statistics over the rows a lane holds, summed over the wave.
__device__ inline void pair_dots(const double2 (&x)[kRowsPerLane], const double2 (&y)[kRowsPerLane], double& al,
                                 double& be, double& gr, double& gi) {
    al = be = gr = gi = 0.0;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const double2 u = x[q], v = y[q];
        al += u.x * u.x + u.y * u.y;
        be += v.x * v.x + v.y * v.y;
        gr += u.x * v.x + u.y * v.y;  // conj(u) v
        gi += u.x * v.y - u.y * v.x;
    }
    al = wave_sum(al);
    be = wave_sum(be);
    gr = wave_sum(gr);
    gi = wave_sum(gi);
}

// Columns whose norm^2 is at most floor2 hold rounding noise (theta = A C has
// rank <= the inner bond, so up to half its columns converge to noise, whose
// direction no relative test can settle): pairs with such a column are not
// rotated. floor2 = min(kNoise^2 |theta|_F^2, kNoiseAbs^2): a skipped column
// keeps S <= 5e-15 and so always falls under the keep rule's S < 1e-14 cut (a
// kept column that was never orthogonalised would be an O(1) error once
// normalised, which is why the floor is capped in absolute terms).
constexpr double kNoise = 7.105427357601002e-15;  // 2^-47
constexpr double kNoiseAbs = 5e-15;
__device__ inline double noise_floor2(double fro2) { return fmin(kNoise * kNoise * fro2, kNoiseAbs * kNoiseAbs); }

// One Jacobi rotation of columns x, y (lane-held rows); returns true if applied.
// `big` is raised when the pair's coupling |x^H y| / (|x| |y|) exceeds big_tol.
__device__ inline bool pair_rotate(double2 (&x)[kRowsPerLane], double2 (&y)[kRowsPerLane], int m, double floor2,
                                   double big_tol = 0.0, bool* big = nullptr) {
    double al, be, gr, gi;
    pair_dots(x, y, al, be, gr, gi);
    if (al <= floor2 || be <= floor2) return false;
    const double gabs = sqrt(gr * gr + gi * gi);
    const double scale = sqrt(al * be);
    if (!(gabs > 0.0) || gabs <= jacobi_tol(m) * scale) return false;
    if (big && gabs > big_tol * scale) *big = true;
    const double2 ph = make_double2(gr / gabs, -gi / gabs);  // conj(gamma)/|gamma|
    const double zeta = (be - al) / (2.0 * gabs);
    const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const double2 u = x[q], v = cmul(y[q], ph);
        x[q] = make_double2(cs * u.x - sn * v.x, cs * u.y - sn * v.y);
        y[q] = make_double2(sn * u.x + cs * v.x, sn * u.y + cs * v.y);
    }
    return true;
}

template <bool XL = false>
__device__ inline void load_col(double2 (&x)[kRowsPerLane], const double2* src, int m, int lane) {
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const int row = lane + 64 * q;
        x[q] = row < m ? gld<XL>(src + row) : make_double2(0, 0);
    }
}
__device__ inline void store_col(const double2 (&x)[kRowsPerLane], double2* dst, int m, int lane) {
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const int row = lane + 64 * q;
        if (row < m) dst[row] = x[q];
    }
}

__global__ void __launch_bounds__(kSvdThreads) k_svd(double2* __restrict__ Wg, int ld, int* dims, int p, int chi,
                                                     double* __restrict__ sv_out, int* __restrict__ perm_out,
                                                     double* trunc, int use_lds, unsigned long long* stats) {
    extern __shared__ double2 lds[];
    __shared__ double s_nrm[kMaxCols];
    __shared__ int s_perm[kMaxCols];
    __shared__ int s_flag;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int Dl = dims[p], Dr = dims[p + 2];
    const int R = 2 * Dl, C = 2 * Dr;
    const bool tall = C <= R;
    const int m = tall ? R : C, c = tall ? C : R;
    double2* W = Wg;
    int lw = ld;
    if (use_lds) {
        W = lds;
        lw = m;
        for (int idx = tid; idx < m * c; idx += blockDim.x) {
            const int col = idx / m, row = idx % m;
            lds[idx] = Wg[(std::size_t)col * ld + row];
        }
        __syncthreads();
    }
    __shared__ unsigned long long s_vis, s_rot;
    if (tid == 0) {
        s_vis = 0;
        s_rot = 0;
    }
    // |theta|_F^2 for the noise floor of pair_rotate
    __shared__ double s_fro[kSvdThreads / 64];
    {
        double a = 0.0;
        for (int idx = tid; idx < m * c; idx += blockDim.x) {
            const int col = idx / m, row = idx % m;
            const double2 v = W[(std::size_t)col * lw + row];
            a += v.x * v.x + v.y * v.y;
        }
        a = wave_sum(a);
        if (lane == 0) s_fro[wave] = a;
        __syncthreads();
    }
    double fro = 0.0;
    for (int w = 0; w < nwaves; ++w) fro += s_fro[w];
    const double floor2 = noise_floor2(fro);
    const int cp = c + (c & 1);
    int sweeps_done = 0;
    unsigned long long visited = 0, rotated = 0;  // pair visits / rotations (wave-local, lane 0 counts)
    for (int sweep = 0; sweep < 60 && cp >= 2; ++sweep) {
        ++sweeps_done;
        if (tid == 0) s_flag = 0;
        __syncthreads();
        for (int r = 0; r < cp - 1; ++r) {
            for (int k = wave; k < cp / 2; k += nwaves) {
                int i, j;
                rr_pair(r, k, cp, i, j);
                if (i >= c || j >= c) continue;
                double2* x = W + (std::size_t)i * lw;
                double2* y = W + (std::size_t)j * lw;
                double2 xa[kRowsPerLane], ya[kRowsPerLane];
                load_col(xa, x, m, lane);
                load_col(ya, y, m, lane);
                ++visited;
                if (pair_rotate(xa, ya, m, floor2)) {
                    ++rotated;
                    if (lane == 0) s_flag = 1;
                    store_col(xa, x, m, lane);
                    store_col(ya, y, m, lane);
                }
            }
            __syncthreads();
        }
        const int f = s_flag;
        __syncthreads();
        if (!f) break;
    }
    // singular values = column norms (FP64)
    for (int j = wave; j < c; j += nwaves) {
        const double2* x = W + (std::size_t)j * lw;
        double s = 0.0;
        for (int row = lane; row < m; row += 64) s += x[row].x * x[row].x + x[row].y * x[row].y;
        s = wave_sum(s);
        if (lane == 0) s_nrm[j] = sqrt(s);
    }
    __syncthreads();
    if (tid < c) {  // stable descending rank
        const double v = rank_key(s_nrm[tid]);
        int rank = 0;
        for (int q = 0; q < c; ++q) {
            const double u = rank_key(s_nrm[q]);
            rank += (u > v) || (u == v && q < tid);
        }
        s_perm[rank] = tid;
    }
    __syncthreads();
    if (tid == 0) {  // mps.cpp:77-85
        double total = 0.0;
        for (int q = 0; q < c; ++q) {
            const double s = s_nrm[s_perm[q]];
            total += s * s;
        }
        int keep = min(chi, c);
        while (keep > 1 && s_nrm[s_perm[keep - 1]] < 1e-14) --keep;
        double disc = 0.0;
        for (int q = keep; q < c; ++q) {
            const double s = s_nrm[s_perm[q]];
            disc += s * s;
        }
        if (total > 0) trunc[blockIdx.x] = fmax(trunc[blockIdx.x], disc / total);
        dims[p + 1] = keep;
    }
    if (lane == 0) {
        atomicAdd(&s_vis, visited * (unsigned long long)m);  // LDS
        atomicAdd(&s_rot, rotated * (unsigned long long)m);
    }
    __syncthreads();
    if (tid == 0) {  // this workgroup's accounting slot
        unsigned long long* slot = stats + 4 * (std::size_t)blockIdx.x;
        slot[0] += 1;
        slot[1] += (unsigned long long)sweeps_done;
        slot[2] += s_vis;
        slot[3] += s_rot;
    }
    for (int q = tid; q < c; q += blockDim.x) {
        sv_out[q] = s_nrm[s_perm[q]];
        perm_out[q] = s_perm[q];
    }
    if (use_lds) {
        for (int idx = tid; idx < m * c; idx += blockDim.x) {
            const int col = idx / m, row = idx % m;
            Wg[(std::size_t)col * ld + row] = lds[idx];
        }
    }
}

// ---- one-sided block Jacobi SVD over a cooperative grid ---------------------
// Columns are split into blocks of 8; each workgroup holds two blocks in LDS and
// the blocks meet in round-robin tournament order, one grid barrier per outer
// step. At the first outer step of a sweep every block also rotates its own
// column pairs, so each sweep visits every column pair exactly once. In the
// cross phase wave w keeps column w of the first block in registers and meets
// column (w + s) mod 8 of the second block in step s; column norms are carried
// across rotations (|x'|^2 = |x|^2 - t|g|, |y'|^2 = |y|^2 + t|g|), recomputed at
// every load and whenever a norm falls below 1/64 of its previous value, so a
// cross pair costs one two-value wave reduction. Rotation formulas, FP64 angles
// and the convergence test are those of k_svd, including the noise floor below.
constexpr int kBlk = 8;
constexpr int kDistThreads = 512;  // 8 waves: one column pair per wave
constexpr int kStatSlots = 1024;   // per-workgroup accounting slots (largest grid)

// Shared LDS of the block-Jacobi SVD (static part).
// Executor phase profile of workgroup 0, in wall-clock ticks (slot meanings
// at k_tebd_lockstep).
constexpr int kHistSweeps = 24;  // sweep histogram: sweeps past the last fold into it
constexpr int kProfHist = 17;    // first histogram slot
constexpr int kProfStep = kProfHist + 3 * kHistSweeps;  // row-layout step parts (cycles)
// wall ticks: whole rounds, sweep ends, the part of a round before its pair work
// (the producer wait included), the part between the pair work and the publication
// (the write-back's stores completing)
constexpr int kProfRound = kProfStep + 5;
constexpr int kProfSlots = kProfRound + 4;

struct SvdShared {
    unsigned long long tm[kProfSlots];  // phase profile (workgroup 0, thread 0)
    double nrm[kMaxCols];
    int perm[kMaxCols];
    double bn[kBlk];  // norms^2 of the LDS block's columns
    double cn[64];    // svd_small: carried norms^2 of the LDS columns
    double cr[64];    // svd_small: each column's norm^2 when last computed exactly
    int rot;
    unsigned long long vis, rots;
    double2 red[kDistThreads / 64][16];  // LQ: per-wave partial row dots
    double2 hh[2];                        // LQ: tau, 1/(alpha - beta)
    double beta;
};

// |x|^2 over the lane-held rows, summed over the wave.
__device__ inline double col_norm2(const double2 (&x)[kRowsPerLane]) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) a += x[q].x * x[q].x + x[q].y * x[q].y;
    return wave_sum(a);
}

// pair_rotate with carried norms^2 al (x) and be (y): only g = x^H y is
// reduced; the norms are updated (recomputed when they shrink by 64x or more).
__device__ inline bool pair_rotate_carried(double2 (&x)[kRowsPerLane], double2 (&y)[kRowsPerLane], int m,
                                           double& al, double& be, double floor2, double big_tol, bool& big) {
    if (al <= floor2 || be <= floor2) return false;
    double gr = 0.0, gi = 0.0;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const double2 u = x[q], v = y[q];
        gr += u.x * v.x + u.y * v.y;  // conj(u) v
        gi += u.x * v.y - u.y * v.x;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        gr += __shfl_xor(gr, off, 64);
        gi += __shfl_xor(gi, off, 64);
    }
    const double gabs = sqrt(gr * gr + gi * gi);
    const double scale = sqrt(al * be);
    if (!(gabs > 0.0) || gabs <= jacobi_tol(m) * scale) return false;
    if (gabs > big_tol * scale) big = true;
    const double2 ph = make_double2(gr / gabs, -gi / gabs);  // conj(gamma)/|gamma|
    const double zeta = (be - al) / (2.0 * gabs);
    const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
        const double2 u = x[q], v = cmul(y[q], ph);
        x[q] = make_double2(cs * u.x - sn * v.x, cs * u.y - sn * v.y);
        y[q] = make_double2(sn * u.x + cs * v.x, sn * u.y + cs * v.y);
    }
    const double al1 = al - t * gabs, be1 = be + t * gabs;
    al = (al1 < al * (1.0 / 64)) ? col_norm2(x) : al1;
    be = (be1 < be * (1.0 / 64)) ? col_norm2(y) : be1;
    return true;
}

// ---- row layout of a tournament round -------------------------------------
// Thread t of the workgroup holds row t (t < m) of the 16 columns of the
// round's block pair in registers: v[q], q < 8 the columns of block bi, q >= 8
// those of block bj. A step rotates 8 disjoint column pairs: every lane forms
// its row's share of the 8 products x^H y, one transposed wave reduction leaves
// each lane one of the 16 partial sums, the waves' partials meet in LDS behind
// one barrier, and every wave turns the totals into the 8 rotations (one pair
// per lane, computed identically by each wave, so no second barrier) and
// applies them to its own rows. The columns never pass through LDS. Column
// norms^2 are exact at the start of the round and carried through its steps
// (recomputed when one shrinks by 64x or more), as in pair_rotate_carried.
struct RowLds {
    double red[2][16][kDistThreads / 64];  // partial sums [parity][value][wave]
    double rn[kDistThreads / 64][16];      // each wave's copy of the carried norms^2
    double par[kDistThreads / 64][8][4];   // each wave's rotations: cs, sn, conj(g)/|g|
};

// Total over the NW waves of partial value (lane & 15) of buffer p: NW
// contiguous doubles per lane, read together.
template <int NW>
__device__ __forceinline__ double row_total(const RowLds& rl, int p, int lane) {
    const double* r = rl.red[p][lane & 15];
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += r[w];
    return t;
}

// Cross-lane moves of doubles without LDS. swap32: lanes 32-63 of a trade with
// lanes 0-31 of b (v_permlane32_swap); swap16: in each half, lanes 16-31 of a
// trade with lanes 0-15 of b (v_permlane16_swap); dpp_f64<CTRL>: the value of
// the lane the DPP control names (0x140 row_mirror: l ^ 15 within a row of 16,
// 0x141 row_half_mirror: l ^ 7 within 8, 0x4E / 0xB1 quad_perm: l ^ 2 / l ^ 1).
__device__ __forceinline__ void swap32(double& a, double& b) {
    const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const auto rl = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)y, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap((unsigned)(x >> 32), (unsigned)(y >> 32), false, false);
    a = __longlong_as_double(((long long)rh[0] << 32) | rl[0]);
    b = __longlong_as_double(((long long)rh[1] << 32) | rl[1]);
}
__device__ __forceinline__ void swap16(double& a, double& b) {
    const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const auto rl = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)y, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap((unsigned)(x >> 32), (unsigned)(y >> 32), false, false);
    a = __longlong_as_double(((long long)rh[0] << 32) | rl[0]);
    b = __longlong_as_double(((long long)rh[1] << 32) | rl[1]);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sums 16 values over the wave's lanes; lane l returns the total of value
// (l >> 2) & 15. Each level pairs every lane with one in the other half of its
// group (32, 16, 8 and 4 lanes; the partners' sets of summed lanes are
// disjoint), keeps half of its values and adds the partner's copy of them;
// the last two levels sum the quad.
__device__ __forceinline__ double wave_reduce16(const double (&d)[16], int lane) {
    const bool b3 = lane & 8, b2 = lane & 4;
    double e[8], f[4], g[2];
#pragma unroll
    for (int i = 0; i < 8; ++i) {  // lanes < 32 keep values 0-7, the others 8-15
        double a = d[i], b = d[i + 8];
        swap32(a, b);
        e[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double a = e[i], b = e[i + 4];
        swap16(a, b);
        f[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) g[i] = (b3 ? f[i + 2] : f[i]) + dpp_f64<0x140>(b3 ? f[i] : f[i + 2]);
    double h = (b2 ? g[1] : g[0]) + dpp_f64<0x141>(b2 ? g[0] : g[1]);
    h += dpp_f64<0x4E>(h);
    h += dpp_f64<0xB1>(h);
    return h;
}

// 1/sqrt(x) and 1/x for finite x > 0: the hardware estimates (v_rsq_f64,
// v_rcp_f64) refined by two Newton steps, y += y (1/2 - x y^2 / 2) and
// y += y (1 - x y); within one unit in the last place over 1e-35..1e35.
__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
#pragma unroll
    for (int i = 0; i < 2; ++i) y = fma(y, fma(-hx * y, y, 0.5), y);
    return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < 2; ++i) y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

// Column pair k of step s: in-block steps (KIND 0, s < 7) pair the columns of
// each block in the rr_pair order of 8 (pairs 0-3 block bi, 4-7 block bj);
// cross steps (KIND 1, s < 8) pair column k of bi with column (k + s) & 7 of bj.
__host__ __device__ constexpr int row_pair_a(int kind, int s, int k) {
    return kind ? k : (k & 3) == 0 ? s + (k & 4) * 2 : (s + (k & 3)) % 7 + (k & 4) * 2;
}
__host__ __device__ constexpr int row_pair_b(int kind, int s, int k) {
    return kind ? 8 + ((k + s) & 7) : (k & 3) == 0 ? 7 + (k & 4) * 2 : (s - (k & 3) + 7) % 7 + (k & 4) * 2;
}

// Raw buffer stores (__builtin_amdgcn_raw_buffer_store_b128): the fourth word of
// the buffer resource (32-bit data format, no swizzle) and the cache-policy bit
// that writes a store through the L2 (sc1).
typedef int v4i32 __attribute__((ext_vector_type(4)));
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr int kCpolSc1 = 16;

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exact norms^2 of the 16 columns into every wave's rn (one reduction). RPT:
// rows per thread (v[r] holds row tid + r * blockDim.x), so 8 / RPT waves.
template <int RPT>
__device__ __forceinline__ void row_norms(const double2 (&v)[RPT][16], RowLds& rl, int& ph, int lane, int wave) {
    double d[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        d[q] = cnorm(v[0][q]);
#pragma unroll
        for (int r = 1; r < RPT; ++r) d[q] += cnorm(v[r][q]);
    }
    const double h = wave_reduce16(d, lane);
    const int p = ph++ & 1;
    if ((lane & 3) == 0) rl.red[p][lane >> 2][wave] = h;
    __syncthreads();
    const double t = row_total<8 / RPT>(rl, p, lane);
    if (lane < 16) rl.rn[wave][lane] = t;
    wave_lds_fence();
}

// The cycle profile of the row-layout steps costs registers in the step's hot
// path, so it exists only in builds with -DTESSERA_STEP_PROFILE=1.
#ifndef TESSERA_STEP_PROFILE
#define TESSERA_STEP_PROFILE 0
#endif
constexpr bool kStepProfile = TESSERA_STEP_PROFILE != 0;

// One step of 8 pair rotations (KIND, S as in row_pair_a). cvalid: bit q set
// when column q exists. Accumulates visited / rotated pairs and the rotation
// flags (bit 0 any, bit 1 a coupling above big_tol).
template <int KIND, int S, int RPT>
__device__ __forceinline__ void row_step(double2 (&v)[RPT][16], RowLds& rl, int& ph, int lane, int wave, unsigned cvalid,
                                int m, double floor2, double big_tol, int& flags, int& nvis, int& nrot,
                                unsigned long long* tm) {
    // tm (thread 0 of the profiled workgroup, in builds with TESSERA_STEP_PROFILE
    // set): cycles of the step's parts
    unsigned long long tc = (kStepProfile && tm) ? __builtin_readcyclecounter() : 0;
    auto mark = [&](int q) {
        if (kStepProfile && tm) {
            const unsigned long long t = __builtin_readcyclecounter();
            tm[q] += t - tc;
            tc = t;
        }
    };
    double d[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double2 x = v[0][row_pair_a(KIND, S, k)], y = v[0][row_pair_b(KIND, S, k)];
        d[2 * k] = x.x * y.x + x.y * y.y;  // conj(x) y
        d[2 * k + 1] = x.x * y.y - x.y * y.x;
#pragma unroll
        for (int r = 1; r < RPT; ++r) {  // the thread's further rows
            const double2 xr = v[r][row_pair_a(KIND, S, k)], yr = v[r][row_pair_b(KIND, S, k)];
            const double re = xr.x * yr.x + xr.y * yr.y, im = xr.x * yr.y - xr.y * yr.x;
            d[2 * k] += re;
            d[2 * k + 1] += im;
        }
    }
    const double h = wave_reduce16(d, lane);
    const int p = ph++ & 1;
    if ((lane & 3) == 0) rl.red[p][lane >> 2][wave] = h;
    mark(0);
    __syncthreads();
    mark(1);
    // pair k = (lane >> 1) & 7: the even lane of each lane pair holds Re x^H y
    // and takes Im from its odd neighbour (every 16 lanes compute the same 8
    // rotations; only the even lanes' results are used)
    const int k = (lane >> 1) & 7, a = row_pair_a(KIND, S, k), b = row_pair_b(KIND, S, k);
    const double gr = row_total<8 / RPT>(rl, p, lane), gi = dpp_f64<0xB1>(gr);
    const bool vis = ((cvalid >> a) & 1) && ((cvalid >> b) & 1);
    double al = rl.rn[wave][a], be = rl.rn[wave][b];
    double cs = 1.0, sn = 0.0, phx = 1.0, phy = 0.0;
    bool rot = false, big = false, need = false;
    if (vis && al > floor2 && be > floor2) {
        // the rotation of pair_rotate_carried with squared tests and refined
        // hardware reciprocal square roots in place of divides and square roots.
        // With d = beta - alpha and zeta = d / (2 |gamma|), the rotation with
        // t = sgn(zeta) / (|zeta| + sqrt(1 + zeta^2)) has
        //   cs^2 = u = (1 + |d| h) / 2,  sn = sgn(d) |gamma| h / sqrt(u),
        //   t |gamma| = sgn(d) |gamma|^2 h / u,  h = 1 / sqrt(4 |gamma|^2 + d^2),
        // so two independent reciprocal square roots (of |gamma|^2 for the phase,
        // and of 4 |gamma|^2 + d^2) and one of u form it; cs^2 + sn^2 = 1 and the
        // phase has unit modulus to rounding, so the transformation stays unitary
        const double g2 = gr * gr + gi * gi, ab = al * be, tol = jacobi_tol(m);
        if (g2 > 0.0 && g2 > tol * tol * ab) {
            rot = true;
            big = g2 > big_tol * big_tol * ab;
            const double d = be - al;
            const double rg = rsq_nr(g2), h = rsq_nr(fma(4.0, g2, d * d));
            phx = gr * rg;  // conj(gamma)/|gamma|
            phy = -gi * rg;
            const double u = fma(0.5 * fabs(d), h, 0.5), q = rsq_nr(u);
            const double sg = d >= 0 ? 1.0 : -1.0;
            cs = u * q;
            sn = sg * (g2 * rg) * h * q;
            const double tg = sg * g2 * h * (q * q);  // t |gamma|
            const double al1 = al - tg, be1 = be + tg;
            need = al1 < al * (1.0 / 64) || be1 < be * (1.0 / 64);
            al = al1;
            be = be1;
        }
    }
    mark(2);
    // unrotated pairs publish the identity (cs 1, sn 0, phase 1), which the
    // application below reproduces exactly
    const bool writer = lane < 16 && !(lane & 1);
    if (writer) {
        rl.par[wave][k][0] = cs;
        rl.par[wave][k][1] = sn;
        rl.par[wave][k][2] = phx;
        rl.par[wave][k][3] = phy;
        if (rot) {
            rl.rn[wave][a] = al;
            rl.rn[wave][b] = be;
        }
    }
    const unsigned long long even = 0x5555ull;  // the writers' lanes
    const unsigned long long rotm = __ballot(rot) & even, bigm = __ballot(big) & even;
    const unsigned long long needm = __ballot(need) & even, vism = __ballot(vis) & even;
    wave_lds_fence();
    mark(3);
    {  // unconditional: a branch around it costs register copies at the join
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double c0 = rl.par[wave][q][0], s0 = rl.par[wave][q][1];
            const double2 e = make_double2(rl.par[wave][q][2], rl.par[wave][q][3]);
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const double2 x = v[r][row_pair_a(KIND, S, q)], y = cmul(v[r][row_pair_b(KIND, S, q)], e);
                v[r][row_pair_a(KIND, S, q)] = make_double2(c0 * x.x - s0 * y.x, c0 * x.y - s0 * y.y);
                v[r][row_pair_b(KIND, S, q)] = make_double2(s0 * x.x + c0 * y.x, s0 * x.y + c0 * y.y);
            }
        }
    }
    mark(4);
    nvis += __popcll(vism);
    nrot += __popcll(rotm);
    flags |= (rotm ? 1 : 0) | (bigm ? 2 : 0);
    if (needm) row_norms<RPT>(v, rl, ph, lane, wave);  // identical in every wave: a uniform branch
}

template <int KIND, int S, int N, int RPT>
__device__ __forceinline__ void row_steps(double2 (&v)[RPT][16], RowLds& rl, int& ph, int lane, int wave,
                                 unsigned cvalid, int m, double floor2, double big_tol, int& flags, int& nvis,
                                 int& nrot, unsigned long long* tm) {
    if constexpr (S < N) {
        row_step<KIND, S, RPT>(v, rl, ph, lane, wave, cvalid, m, floor2, big_tol, flags, nvis, nrot, tm);
        row_steps<KIND, S + 1, N, RPT>(v, rl, ph, lane, wave, cvalid, m, floor2, big_tol, flags, nvis, nrot, tm);
    }
}

// A whole tournament round of block pair (bi, bj) in the row layout: load the
// 16 columns into the caller's registers v (block bi only when load_a; else v
// still holds it from the previous round), the in-block pairs (once per sweep),
// the cross pairs, write back (block bi only when store_a). Returns the
// rotation flags; counts visited and rotated pairs. `lap` marks the profile phases (12 load, 13 in-block, 14 cross,
// 15 write-back). ph counts the LDS reductions of the workgroup: consecutive
// reductions alternate between the two partial buffers, so one barrier each
// suffices (a wave can be at most one reduction ahead of another).
template <bool XL, int RPT, class Lap>
__device__ __forceinline__ int round_rows(double2* __restrict__ Wg, int ld, int m, int c, int bi, int bj, bool inblock,
                                 double floor2, double big_tol, RowLds& rl, int& ph, int& nvis, int& nrot,
                                 unsigned long long* tm, double2 (&v)[RPT][16], bool load_a, bool store_a,
                                 Lap&& lap) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int kThreads = kDistThreads / RPT;  // row tid + r * kThreads in v[r]
    // the write-back's buffer resource: Wg's ld x ld matrix (Wg and ld are uniform;
    // read from lane 0 so that the resource sits in scalar registers)
    const unsigned long long wa = reinterpret_cast<unsigned long long>(Wg);
    const unsigned long long wu = ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(wa >> 32)) << 32) |
                                  (unsigned)__builtin_amdgcn_readfirstlane((int)wa);
    const int wbytes = __builtin_amdgcn_readfirstlane((int)((std::size_t)ld * ld * sizeof(double2)));
    const __amdgpu_buffer_rsrc_t wrsrc =
        __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(wu), 0, wbytes, kBufferRsrcWord3);
    unsigned cvalid = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = (q < kBlk ? bi : bj) * kBlk + (q & (kBlk - 1));
        const bool ok = col < c;
        cvalid |= ok ? 1u << q : 0u;
        if (q >= kBlk || load_a)
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int row = tid + r * kThreads;
                v[r][q] = (ok && row < m) ? gld<XL>(Wg + (std::size_t)col * ld + row) : make_double2(0.0, 0.0);
            }
    }
    int flags = 0;
    row_norms<RPT>(v, rl, ph, lane, wave);
    lap(12);
    if (inblock)
        row_steps<0, 0, kBlk - 1, RPT>(v, rl, ph, lane, wave, cvalid, m, floor2, big_tol, flags, nvis, nrot, tm);
    lap(13);
    row_steps<1, 0, kBlk, RPT>(v, rl, ph, lane, wave, cvalid, m, floor2, big_tol, flags, nvis, nrot, tm);
    lap(14);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = (q < kBlk ? bi : bj) * kBlk + (q & (kBlk - 1));
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int row = tid + r * kThreads;
            if ((q >= kBlk || store_a) && ((cvalid >> q) & 1) && row < m) {
                if constexpr (XL) {
                    Wg[(std::size_t)col * ld + row] = v[r][q];
                } else {
                    // 16-byte stores with sc1 write through the L2, so the round's
                    // release finds no dirty lines of the block to write back
                    v4i32 d;
                    __builtin_memcpy(&d, &v[r][q], sizeof(d));
                    __builtin_amdgcn_raw_buffer_store_b128(d, wrsrc, (int)(((std::size_t)col * ld + row) * sizeof(double2)),
                                                           0, kCpolSc1);
                }
            }
        }
    }
    lap(15);
    return flags;
}

// ---- LQ preconditioning of the block-Jacobi SVD ----------------------------
// W (m x c, m >= c, column-major) <- W Q = L, lower trapezoidal, by Householder
// reflectors applied from the right (reflector k zeroes row k beyond column k).
// L has the left singular vectors and singular values of W, and the one-sided
// Jacobi converges on L's columns in about half the sweeps it needs on W's
// (the rows of a QR/LQ factor are graded; on the thetas of a saturated run 10-14
// sweeps instead of 22-43 in a CPU study). Rows go in tiles of 16; tile t
// belongs to worker t mod nworkers, and inside a workgroup each tile row belongs
// to one 32-lane team that holds it in registers (lane l: columns r0 + l + 32 u).
// Panel q (rows 16q..16q+15, columns r0 = 16q onward) is factorised by its
// owner: the team of row i builds reflector i from its own row and writes it to
// LDS, one workgroup barrier publishes it, and every later row's team applies
// it to its row (a half-wave reduction, no barrier). Reflector k (v_k = 1,
// entries at columns >= k) is kept as column k of `aux` with tau in taus[k].
// After one group barrier per panel every worker stages the panel's reflectors
// in LDS once and each team applies all of them to its row of every later tile
// it owns, without further barriers.
// Reflectors in LDS: entry jj of reflector i at vl[i * S + jj], S = c + 1.
constexpr int kLqCols = kMaxCols / 32;  // register slots of a team lane (columns per lane)

__host__ __device__ constexpr std::size_t lq_lds_bytes(int c) {
    return (std::size_t)16 * (c + 1) * sizeof(double2) + 16 * sizeof(double2);
}

// sum over the 32 lanes of this lane's half of the wave: the mirror levels leave
// every lane of a row of 16 with the row's total, the permlane swap adds the
// other row's
__device__ __forceinline__ double half_sum(double v) {
    v += dpp_f64<0x140>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0xB1>(v);
    double a = v, b = v;
    swap16(a, b);
    return v + ((threadIdx.x & 16) ? a : b);
}

// row `row` of W, columns r0 + l32 + 32 u (u < kLqCols, inside r0..r0+n-1) into x
template <bool XL>
__device__ __forceinline__ void lq_load_row(const double2* __restrict__ Wg, int ld, int row, int r0, int n, int m,
                                            int l32, double2 (&x)[kLqCols]) {
#pragma unroll
    for (int u = 0; u < kLqCols; ++u) {
        const int jj = l32 + 32 * u;
        x[u] = (jj < n && row < m) ? gld<XL>(Wg + (std::size_t)(r0 + jj) * ld + row) : make_double2(0.0, 0.0);
    }
}
__device__ __forceinline__ void lq_store_row(double2* __restrict__ Wg, int ld, int row, int r0, int n, int m,
                                             int l32, const double2 (&x)[kLqCols]) {
#pragma unroll
    for (int u = 0; u < kLqCols; ++u) {
        const int jj = l32 + 32 * u;
        if (jj < n && row < m) Wg[(std::size_t)(r0 + jj) * ld + row] = x[u];
    }
}
// x <- x (I - tau v v^H) for reflector v at vl (entries from column i on)
__device__ __forceinline__ void lq_reflect(double2 (&x)[kLqCols], const double2* vl, int i, int n, int l32,
                                           double2 tau) {
    double2 w = make_double2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < kLqCols; ++u) {
        const int jj = l32 + 32 * u;
        if (jj >= i && jj < n) w = cadd(w, cmul(x[u], vl[jj]));
    }
    w.x = half_sum(w.x);
    w.y = half_sum(w.y);
    const double2 tw = cmul(tau, w);
#pragma unroll
    for (int u = 0; u < kLqCols; ++u) {
        const int jj = l32 + 32 * u;
        if (jj >= i && jj < n) x[u] = csub(x[u], cmulc(vl[jj], tw));
    }
}

template <class Sync, bool XL, int RPT = 1>
__device__ void lq_precondition(double2* __restrict__ Wg, int ld, int m, int c, double2* __restrict__ aux,
                                double2* __restrict__ taus, int me, int nworkers, Sync& sync, double2* lds,
                                int panels, bool active) {
    // RPT tile rows per team: 16 / RPT teams of 32 lanes (blockDim.x = 512 / RPT);
    // tile row i belongs to team i % kTeams, register slot i / kTeams
    constexpr int kTeams = 16 / RPT;
    const int tid = threadIdx.x, team = tid >> 5, l32 = tid & 31;
    const int S = c + 1;
    double2* vl = lds;
    double2* tauv = lds + (std::size_t)16 * S;  // taus of the panel in LDS
    const int ntiles = (m + 15) / 16;
    const double2 z = make_double2(0.0, 0.0);
    double2 x[RPT][kLqCols];
    for (int q = 0; q < panels; ++q) {
        const int r0 = q * 16;
        const bool live = active && r0 < c;
        const int n = c - r0, nb = min(16, c - r0);
        const bool owner = live && q % nworkers == me;
        if (owner) {  // factorise panel q
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) lq_load_row<XL>(Wg, ld, q * 16 + team + kTeams * sl, r0, n, m, l32, x[sl]);
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl)
                for (int i = sl * kTeams; i < min(nb, (sl + 1) * kTeams); ++i) {
                    if (team == i - sl * kTeams) {
                        // reflector i from row i beyond column i: row (I - tau v v^H) = (beta, 0, ...)
                        double2 (&xr)[kLqCols] = x[sl];
                        double part = 0.0;
#pragma unroll
                        for (int u = 0; u < kLqCols; ++u) {
                            const int jj = l32 + 32 * u;
                            if (jj > i && jj < n) part += cnorm(xr[u]);
                        }
                        const double xn2 = half_sum(part);
                        // column i < 16 sits in slot 0 of lane i
                        const double2 alpha = make_double2(__shfl(xr[0].x, i, 32), -__shfl(xr[0].y, i, 32));
                        double2 tau, scale;
                        double beta;
                        if (xn2 == 0.0 && alpha.y == 0.0) {
                            tau = z;
                            scale = z;
                            beta = alpha.x;
                        } else {
                            beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
                            const double rb = 1.0 / beta;
                            tau = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
                            const double2 d = make_double2(alpha.x - beta, alpha.y);  // 1 / (alpha - beta)
                            const double rdd = 1.0 / (d.x * d.x + d.y * d.y);
                            scale = make_double2(d.x * rdd, -d.y * rdd);
                        }
                        // into LDS only: a global store here would hold up the barrier
                        // below (it waits for the workgroup's outstanding stores)
                        double2* vi = vl + i * S;
#pragma unroll
                        for (int u = 0; u < kLqCols; ++u) {
                            const int jj = l32 + 32 * u;
                            if (jj >= i && jj < n) {
                                const double2 e = jj == i ? make_double2(1.0, 0.0) : cmul(cconj(xr[u]), scale);
                                vi[jj] = e;
                                xr[u] = jj == i ? make_double2(beta, 0.0) : z;  // row i of L
                            }
                        }
                        if (l32 == 0) tauv[i] = tau;
                    }
                    __syncthreads();
#pragma unroll
                    for (int s2 = sl; s2 < RPT; ++s2)
                        if (team + kTeams * s2 > i) lq_reflect(x[s2], vl + i * S, i, n, l32, tauv[i]);
                }
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) lq_store_row(Wg, ld, q * 16 + team + kTeams * sl, r0, n, m, l32, x[sl]);
            // the panel's reflectors (LDS, complete after the last barrier above) and
            // taus to aux and taus for the other workers
            for (int idx = tid; idx < nb * n; idx += (int)blockDim.x) {
                const int k = idx / n, jj = idx - k * n;
                if (jj >= k) aux[(std::size_t)(r0 + k) * ld + r0 + jj] = vl[k * S + jj];
            }
            if (tid < nb) taus[r0 + tid] = tauv[tid];
        }
        sync();
        if (!live) continue;
        bool staged = owner;  // the owner's LDS already holds the panel
        for (int t = q + 1; t < ntiles; ++t) {
            if (t % nworkers != me) continue;
            if (!staged) {  // the panel's reflectors into LDS, all loads of a thread in flight
                for (int base = tid; base < nb * n; base += 4 * (int)blockDim.x) {
                    double2 v[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const int idx = base + h * (int)blockDim.x, k = idx / n, jj = idx - k * n;
                        v[h] = (idx < nb * n && jj >= k) ? gld<XL>(aux + (std::size_t)(r0 + k) * ld + r0 + jj) : z;
                    }
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const int idx = base + h * (int)blockDim.x, k = idx / n, jj = idx - k * n;
                        if (idx < nb * n) vl[k * S + jj] = v[h];
                    }
                }
                if (tid < nb) tauv[tid] = gld<XL>(taus + r0 + tid);
                __syncthreads();
                staged = true;
            }
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) lq_load_row<XL>(Wg, ld, t * 16 + team + kTeams * sl, r0, n, m, l32, x[sl]);
            for (int i = 0; i < nb; ++i)
#pragma unroll
                for (int sl = 0; sl < RPT; ++sl) lq_reflect(x[sl], vl + i * S, i, n, l32, tauv[i]);
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) lq_store_row(Wg, ld, t * 16 + team + kTeams * sl, r0, n, m, l32, x[sl]);
        }
        __syncthreads();  // the next panel rewrites the LDS reflectors
    }
    sync();
}



// ---- convergence check on the matrix cores --------------------------------
// After a sweep with few rotations: every column pair's coupling from one Gram
// matrix G = W^H W (16 x 16 tiles of complex FP64 MFMA products, one tile per
// wave at a time, K = the rows in steps of 4) against the Jacobi tolerance,
// with the sweep's noise floor and exact column norms: the test a sweep applies
// pair by pair, here to every pair at once. When no pair fails, the SVD has
// converged and ends without the sweep that would only confirm it (that sweep's
// own products could round a coupling at the tolerance to the other side; the
// decision is the check's in every layout, so the layouts stay identical).
// Every worker takes every (nworkers * 8)-th tile of the upper triangle; the
// workers' results meet in *vflag (zeroed again by worker 0 after the last
// read). compute = false joins only the barriers.
template <class Sync, bool XL>
__device__ bool gram_converged(const double2* __restrict__ Wg, int ld, int m, int c, double floor2,
                               double* __restrict__ nrm, int* vflag, int me, int nworkers, Sync& sync,
                               bool compute) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wpg = (int)(blockDim.x >> 6);
    const int nw = nworkers * wpg, gw = me * wpg + wave;
    if (!compute) c = 0;  // only the barriers (lockstep: another group checks)
    for (int col = gw; col < c; col += nw) {  // exact norms^2
        const double2* x = Wg + (std::size_t)col * ld;
        double a = 0.0;
        for (int row = lane; row < m; row += 64) a += cnorm(gld<XL>(x + row));
        a = wave_sum(a);
        if (lane == 0) nrm[col] = a;
    }
    sync();
    const int nt = (c + 15) / 16, ntt = nt * (nt + 1) / 2;
    const double tol = jacobi_tol(m), tol2 = tol * tol;
    const int lg = lane >> 4, lc = lane & 15;
    const double2 z = make_double2(0.0, 0.0);
    bool bad = false;
    for (int tt = gw; tt < ntt; tt += nw) {
        int I = 0, rem = tt;  // tile (I, J), I <= J, row-major over the upper triangle
        while (rem >= nt - I) {
            rem -= nt - I;
            ++I;
        }
        const int J = I + rem;
        const int ci = I * 16 + lc, cj = J * 16 + lc;  // this lane's operand columns
        const double2* pa = Wg + (std::size_t)min(ci, c - 1) * ld;
        const double2* pb = Wg + (std::size_t)min(cj, c - 1) * ld;
        CTile g;
        g.zero();
        for (int k0 = 0; k0 < m; k0 += 32) {  // eight K steps of loads in flight
            double2 a[8], b[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + 4 * u + lg;
                a[u] = (k < m && ci < c) ? cconj(gld<XL>(pa + k)) : z;
                b[u] = (k < m && cj < c) ? gld<XL>(pb + k) : z;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) g.mac(a[u], b[u]);
        }
        // this lane holds G[16 I + lg + 4 e][16 J + lc]
        const int j = J * 16 + lc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = I * 16 + lg + 4 * e;
            if (i < j && j < c) {
                const double al = gld<XL>(nrm + i), be = gld<XL>(nrm + j);
                const double2 gg = g.get(e);
                const double g2 = gg.x * gg.x + gg.y * gg.y;
                if (al > floor2 && be > floor2 && g2 > 0.0 && g2 > tol2 * al * be) bad = true;
            }
        }
    }
    if (__ballot(bad) && lane == 0) atomicOr(vflag, 1);
    sync();
    const bool ok = compute && __hip_atomic_load(vflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    sync();
    if (me == 0 && tid == 0) *vflag = 0;
    return ok;
}

// How a block-Jacobi SVD is synchronised with others in the same launch. Alone:
// every sync() is its own, rounds and sweeps follow its own matrix. Lockstep:
// several SVDs (one per group of workgroups) share one grid barrier, so every
// group runs `rounds` outer steps per sweep and the sweep loop ends when no
// group rotated (`gflags`, cleared by the grid leader); a group whose own sweep
// rotated nothing stops working but keeps joining the barriers.
struct SvdMode {
    int rounds = -1;            // < 0: own tournament length
    int lq_panels = -1;         // LQ panels (< 0: own ceil(c/16))
    int* gflags = nullptr;      // grid-wide rotation flags [2], a check asked for [1]
    bool grid_leader = false;   // clears gflags
    bool active = true;         // false: this group has no matrix
    bool prof = false;          // this workgroup records the phase profile in sh.tm
    // point-to-point rounds (one block pair per worker): per-worker round counters,
    // [nworkers] x 64 ints; null: a full barrier after every round
    int* epochs = nullptr;
    int* err = nullptr;         // raised by a round wait that gave up
    long long spin_limit = 0;   // polls before a round wait gives up
    int naps = 1;               // 64-cycle sleeps between polls
    // > 0: the sweeps end after a sweep whose rotations all had a relative coupling
    // |x^H y| / (|x| |y|) of at most stop_ratio (one-sided Jacobi converges
    // quadratically, so the next sweep's couplings are O(stop_ratio^2) and below
    // the rotation threshold); 0: they end after a sweep that rotates nothing
    double stop_ratio = 0.0;
    bool rows = true;           // rounds in the row layout (round_rows); false: columns through LDS
    // a sweep with few rotations is followed by a check of every pair on the matrix
    // cores (gram_converged) that ends the SVD instead of a sweep that rotates nothing
    bool gram = true;
    bool step_prof = false;     // with prof: cycle profile of the row-layout steps (slows them)
};

// Worker holding block b in tournament round r of rr_pair with one pair per
// worker: block cp - 1 and block r belong to pair 0, any other block to the pair
// k in [1, cp/2 - 1] with b = r + k or b = r - k (mod cp - 1).
__device__ inline int rr_holder(int r, int b, int cp) {
    const int L = cp - 1;
    if (b == L || b == r) return 0;
    const int k1 = ((b - r) % L + L) % L;
    return (k1 >= 1 && k1 <= cp / 2 - 1) ? k1 : ((r - b) % L + L) % L;
}

// Resident-block tournament of 2P blocks on P workers (P a power of two, one
// pair per worker and round; 2P - 1 rounds, every block pair once). Level l
// (s = P >> l rounds) splits the blocks into sets of 2s, served by s workers
// each: worker j of a set keeps block j of the set for the whole level and in
// round t meets block s + (j + t) mod s, which it receives from worker j + 1
// of the set. So inside a level each worker loads, stores and waits for one
// block per round; the other stays in its registers.
__device__ inline void tour_level(int r, int P, int& s, int& t) {
    int base = 0;
    s = P;
    while (s > 1 && r >= base + s) {
        base += s;
        s >>= 1;
    }
    t = r - base;
}
__device__ inline void tour_pair(int r, int w, int P, int& a, int& b) {
    int s, t;
    tour_level(r, P, s, t);
    const int g = w / s, j = w % s;
    a = g * 2 * s + j;
    b = g * 2 * s + s + (j + t) % s;
}
__device__ inline int tour_holder(int r, int blk, int P) {
    int s, t;
    tour_level(r, P, s, t);
    const int g = blk / (2 * s), off = blk % (2 * s);
    return off < s ? g * s + off : g * s + (off - s - t + s) % s;
}

// Thread 0 of a workgroup waits until a round counter reaches `target`; every
// wait ends (counter reached, another wait gave up, or the poll limit, which
// raises *err).
// Between polls the thread sleeps `naps` times 64 cycles. The polls are relaxed
// (a device-scope acquire would invalidate the XCD's L2 lines that other
// workgroups are reading at every poll); one acquire fence follows the poll that
// saw the target, which synchronises with the producer's release store that the
// poll read. The error word is read every 32nd poll only, so a poll costs one
// memory round trip.
template <bool XL = false>
__device__ inline void wait_counter(int* ctr, int target, int* err, long long limit, int naps = 1) {
    for (long long spins = 0; __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
         ++spins) {
        if (spins >= limit ||
            ((spins & 31) == 31 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        for (int q = 0; q < naps; ++q) __builtin_amdgcn_s_sleep(1);
    }
    // XCD-local: nothing to acquire; the caller reads the producer's data with
    // L1-bypassing loads
    if constexpr (!XL) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Block-Jacobi SVD run by `nworkers` workgroups that synchronise with `sync()`.
// Worker `me` takes block pairs me, me + nworkers, ... of each tournament round
// (P = nbp / 2 pairs); workers with me >= P only join the barriers. The leader (me == 0) sorts, truncates and
// writes dims[p+1], sv_out, perm_out; callers must synchronise again before
// reading them. Accounting goes to this workgroup's slots: stats[4*b + 0..3] =
// calls, sweeps, visited and rotated pair-rows, trunc[b] = max discarded weight
// (b = blockIdx.x; the host reduces the slots, no cross-workgroup atomics).
template <class Sync, bool XL = false, int RPT = 1>
__device__ void svd_blocks(double2* __restrict__ Wg, int ld, int* dims, int p, int chi, double* __restrict__ sv_out,
                           int* __restrict__ perm_out, double* trunc, int* flags, unsigned long long* stats, int me,
                           int nworkers, Sync& sync, double2* lds, SvdShared& sh, const SvdMode& mode,
                           double2* aux, double2* taus) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = kDistThreads / RPT / 64;
    const int Dl = mode.active ? dims[p] : 1, Dr = mode.active ? dims[p + 2] : 1;
    const int R = 2 * Dl, C = 2 * Dr;
    const bool tall = C <= R;
    const int m = tall ? R : C, c = tall ? C : R;
    const int nb = (c + kBlk - 1) / kBlk;
    const int nbp = nb + (nb & 1) < 2 ? 2 : nb + (nb & 1);
    const int P = nbp / 2;
    const int rounds = mode.rounds >= 0 ? mode.rounds : nbp - 1;
    if (me == 0 && tid == 0) {
        flags[0] = 0;
        flags[1] = 0;
        for (int q = 4; q < 8; ++q) flags[q] = 0;
    }
    if (tid == 0) {
        sh.vis = 0;
        sh.rots = 0;
    }
    sync();
    // |theta|_F^2 for the noise floor: column norms^2 into sv_out (scratch), then
    // summed in column order by every wave, so every worker and every layout gets
    // the same bits
    if (mode.active)
        for (int col = me * nwaves + wave; col < c; col += nworkers * nwaves) {
            const double2* x = Wg + (std::size_t)col * ld;
            double a = 0.0;
            for (int row = lane; row < m; row += 64) a += cnorm(gld<XL>(x + row));
            a = wave_sum(a);
            if (lane == 0) sv_out[col] = a;
        }
    sync();
    double fro2 = 0.0;
    if (mode.active) {
        double s = 0.0;
        for (int col = lane; col < c; col += 64) s += gld<XL>(sv_out + col);
        fro2 = wave_sum(s);
    }
    const double floor2 = noise_floor2(fro2);
    const bool prof = mode.prof && tid == 0;
    unsigned long long tp = prof ? wall_clock64() : 0;
    if (aux) {  // W <- L of W = L Q^H (same left singular vectors and values)
        const int panels = mode.lq_panels >= 0 ? mode.lq_panels : (c + 15) / 16;
        lq_precondition<Sync, XL, RPT>(Wg, ld, m, c, aux, taus, me, nworkers, sync, lds, panels, mode.active);
    }
    if (prof) {
        const unsigned long long t = wall_clock64();
        sh.tm[1] += t - tp;
        tp = t;
    }
    bool working = mode.active;
    // Point-to-point rounds: inside a sweep, round r + 1 of worker me needs only
    // the two blocks that their holders wrote back in round r, so instead of a
    // barrier each worker publishes a round counter (release) and waits for its
    // two producers' counters (acquire). Each block's home columns are written
    // only by its current holder and read only by the next one. The last round of
    // a sweep keeps the full barrier (the rotation flags are read after it).
    const bool p2p = mode.epochs != nullptr && nworkers == P && mode.active;
    // row layout, at most one pair per worker and P a power of two: the
    // resident-block tournament (tour_pair), else rr_pair. The ordering depends
    // only on the shape, so every synchronisation mode applies the same rotations.
    // (RPT > 1: always the row layout; the LDS-column rounds need 8 waves)
    const bool rows = RPT > 1 || mode.rows;
    const bool tour = rows && nworkers >= P && (P & (P - 1)) == 0;
    double2 rv[RPT][16];  // the row layout's registers, kept across rounds for a resident block
    int ep = p2p ? __hip_atomic_load(&mode.epochs[(std::size_t)me * 64], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    int sweeps_done = 0, ph = 0;
    unsigned long long nvis = 0, nrot = 0;
    const bool gram = mode.gram && rows && mode.stop_ratio == 0.0;
    // a check pays off once a sweep rotates at most about 1/256 of the pairs
    const int gram_rot = max(16, c * (c - 1) / 512);
    for (int sweep = 0; sweep < 40; ++sweep) {
        int srot = 0;    // thread 0: this worker's rotations in the sweep
        int sflags = 0;  // thread 0: its rotation flags not yet published
        if (working) ++sweeps_done;
        if (prof && working) sh.tm[kProfHist + 2 * kHistSweeps + min(sweep, kHistSweeps - 1)] += 1;
        for (int r = 0; r < rounds; ++r) {
            const unsigned long long tround = prof ? wall_clock64() : 0;
            unsigned long long tpost = tround;
            if (p2p && r > 0) {
                const unsigned long long tw = prof ? wall_clock64() : 0;
                if (tid == 0) {
                    int bi, bj, pa, pb;
                    if (tour) {
                        tour_pair(r, me, P, bi, bj);
                        pa = tour_holder(r - 1, bi, P);
                        pb = tour_holder(r - 1, bj, P);
                    } else {
                        rr_pair(r, me, nbp, bi, bj);
                        pa = rr_holder(r - 1, bi, nbp);
                        pb = rr_holder(r - 1, bj, nbp);
                    }
                    // device scope: each wait ends with an acquire fence, so the
                    // producers' blocks (behind their release) are visible to the whole
                    // workgroup after the barrier below; XCD-local: the blocks are read
                    // with L1-bypassing loads
                    if (pa != me)
                        wait_counter<XL>(&mode.epochs[(std::size_t)pa * 64], ep, mode.err, mode.spin_limit, mode.naps);
                    if (pb != me)
                        wait_counter<XL>(&mode.epochs[(std::size_t)pb * 64], ep, mode.err, mode.spin_limit, mode.naps);
                }
                __syncthreads();
                if (prof) sh.tm[8] += wall_clock64() - tw;
            }
            // worker `me` takes pairs me, me + nworkers, ... of the round (one
            // pair when there are P workers, all of them for a lone workgroup)
            if (working && r < nbp - 1)
            for (int pi = me; pi < P; pi += nworkers) {
                int bi, bj;
                bool load_a = true, store_a = true;
                if (tour) {
                    tour_pair(r, pi, P, bi, bj);
                    int pa, pb;
                    if (r > 0) {  // bi stayed in the registers since the last round
                        tour_pair(r - 1, pi, P, pa, pb);
                        load_a = pa != bi;
                    }
                    if (r + 1 < rounds && r + 1 < nbp - 1) {  // bi stays for the next round
                        tour_pair(r + 1, pi, P, pa, pb);
                        store_a = pa != bi;
                    }
                } else {
                    rr_pair(r, pi, nbp, bi, bj);
                }
                if (tid == 0) sh.rot = 0;
                unsigned long long tr0 = prof ? wall_clock64() : 0;
                if (prof) sh.tm[kProfRound + 2] += tr0 - tround;
                auto lap = [&](int slot) {  // round phases of workgroup 0
                    if (prof) {
                        const unsigned long long t = wall_clock64();
                        sh.tm[slot] += t - tr0;
                        tr0 = t;
                    }
                };
                int rflags = 0, wrot = 0;
                tpost = 0;
                if (rows) {
                    int nv = 0, nr = 0;
                    rflags = round_rows<XL, RPT>(Wg, ld, m, c, bi, bj, r == 0, floor2, mode.stop_ratio,
                                            *reinterpret_cast<RowLds*>(lds), ph, nv, nr,
                                            (prof && mode.step_prof) ? sh.tm + kProfStep : nullptr, rv, load_a,
                                            store_a, lap);
                    tpost = prof ? wall_clock64() : 0;
                    if (tid == 0) {  // every wave holds the same counts
                        nvis += (unsigned long long)nv;
                        nrot += (unsigned long long)nr;
                        wrot = nr;
                        srot += nr;
                    }
                } else if constexpr (RPT == 1) {
                double2 xa[kRowsPerLane], ya[kRowsPerLane];
                // LDS column q <-> block (q < 8 ? bi : bj), member q & 7; a column's
                // rows go through registers with all of a lane's loads in flight at
                // once (a load-store loop would wait out each load in turn)
                for (int q = wave; q < 2 * kBlk; q += kDistThreads / 64) {
                    const int col = (q < kBlk ? bi : bj) * kBlk + (q & (kBlk - 1));
                    if (col < c) {
                        load_col<XL>(xa, Wg + (std::size_t)col * ld, m, lane);
                        store_col(xa, lds + (std::size_t)q * m, m, lane);
                    }
                }
                __syncthreads();
                lap(12);
                bool rotated = false, big = false;
                if (r == 0) {  // pairs inside each block, once per sweep
                    for (int s = 0; s < kBlk - 1; ++s) {
                        const int half = wave >> 2, k = wave & 3;  // waves 0-3 block bi, 4-7 block bj
                        int i, j;
                        rr_pair(s, k, kBlk, i, j);
                        const int blk = half ? bj : bi;
                        if (blk * kBlk + i < c && blk * kBlk + j < c) {
                            double2* px = lds + (std::size_t)(half * kBlk + i) * m;
                            double2* py = lds + (std::size_t)(half * kBlk + j) * m;
                            load_col(xa, px, m, lane);
                            load_col(ya, py, m, lane);
                            if (pair_rotate(xa, ya, m, floor2, mode.stop_ratio, &big)) {
                                rotated = true;
                                ++nrot;
                                ++wrot;
                                store_col(xa, px, m, lane);
                                store_col(ya, py, m, lane);
                            }
                            ++nvis;
                        }
                        __syncthreads();
                    }
                }
                lap(13);
                // cross pairs with carried norms
                const bool mine = bi * kBlk + wave < c;
                double al = 0.0;
                if (mine) {
                    load_col(xa, lds + (std::size_t)wave * m, m, lane);
                    al = col_norm2(xa);
                }
                if (bj * kBlk + wave < c) {
                    load_col(ya, lds + (std::size_t)(kBlk + wave) * m, m, lane);
                    const double b2 = col_norm2(ya);
                    if (lane == 0) sh.bn[wave] = b2;
                }
                __syncthreads();
                for (int s = 0; s < kBlk; ++s) {
                    const int j = (wave + s) & (kBlk - 1);
                    if (mine && bj * kBlk + j < c) {
                        double2* py = lds + (std::size_t)(kBlk + j) * m;
                        load_col(ya, py, m, lane);
                        double be = sh.bn[j];
                        if (pair_rotate_carried(xa, ya, m, al, be, floor2, mode.stop_ratio, big)) {
                            rotated = true;
                            ++nrot;
                            ++wrot;
                            store_col(ya, py, m, lane);
                            if (lane == 0) sh.bn[j] = be;
                        }
                        ++nvis;
                    }
                    __syncthreads();
                }
                if (mine) store_col(xa, lds + (std::size_t)wave * m, m, lane);
                if (rotated && lane == 0) atomicOr(&sh.rot, big ? 3 : 1);
                __syncthreads();
                lap(14);
                for (int q = wave; q < 2 * kBlk; q += kDistThreads / 64) {
                    const int col = (q < kBlk ? bi : bj) * kBlk + (q & (kBlk - 1));
                    if (col < c) {
                        load_col(xa, lds + (std::size_t)q * m, m, lane);
                        store_col(xa, Wg + (std::size_t)col * ld, m, lane);
                    }
                }
                __syncthreads();  // the next pair of this worker reuses the LDS columns
                lap(15);
                rflags = sh.rot;
                }
                if (mode.prof && lane == 0 && wrot)
                    atomicAdd(&sh.tm[kProfHist + min(sweep, kHistSweeps - 1)], (unsigned long long)wrot);
                if (tid == 0) {
                    sflags |= rflags;
                    if (mode.prof && (rflags & 2))
                        sh.tm[kProfHist + kHistSweeps + min(sweep, kHistSweeps - 1)] += 1;
                }
            }
            if (p2p && r < rounds - 1) {
                const unsigned long long tpub = prof ? wall_clock64() : 0;
                if (prof) sh.tm[kProfRound + 3] += tpub - tpost;
                // publish this round's blocks: every wave waits until its write-back
                // stores are complete in the L2 (s_waitcnt of all counters; stores
                // count in vmcnt on CDNA), the workgroup barrier follows, and thread 0's
                // agent-scope release store then writes the L2 back (the whole XCD's
                // dirty lines, these stores among them) before it stores the counter;
                // no further fence (an acquire here would only invalidate the L2)
                ++ep;
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                if (tid == 0) {
                    if constexpr (XL) {  // the stores are complete in the shared L2
                        __hip_atomic_store(&mode.epochs[(std::size_t)me * 64], ep, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    } else {
                        __hip_atomic_store(&mode.epochs[(std::size_t)me * 64], ep, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                if (prof) {
                    sh.tm[9] += 1;
                    sh.tm[16] += wall_clock64() - tpub;
                }
            } else {
                const unsigned long long tw = prof ? wall_clock64() : 0;
                // the sweep's rotation flags, read only after its last barrier
                if (tid == 0 && sflags) {
                    atomicOr(&flags[sweep & 1], sflags);
                    if (mode.gflags) atomicOr(&mode.gflags[sweep & 1], sflags);
                    sflags = 0;
                }
                if (gram && r == rounds - 1 && tid == 0 && srot) atomicAdd(&flags[4 + (sweep & 1)], srot);
                sync();
                if (prof) {
                    sh.tm[8] += wall_clock64() - tw;
                    sh.tm[9] += 1;
                }
            }
            if (prof) sh.tm[kProfRound] += wall_clock64() - tround;
        }
        const unsigned long long tend = prof ? wall_clock64() : 0;
        const int f = __hip_atomic_load(&flags[sweep & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int gf = mode.gflags ? __hip_atomic_load(&mode.gflags[sweep & 1], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                                   : f;
        const int cnt = __hip_atomic_load(&flags[4 + (sweep & 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // this SVD asks for the convergence check; in lockstep every group joins the
        // check's grid barriers when any group asks (gflags[2])
        const bool want = gram && working && (f & 2) && cnt <= gram_rot;
        if (want && mode.gflags && me == 0 && tid == 0) atomicOr(&mode.gflags[2], 1);
        sync();
        if (me == 0 && tid == 0) {
            flags[sweep & 1] = 0;
            flags[4 + (sweep & 1)] = 0;
        }
        if (mode.grid_leader && tid == 0 && mode.gflags) mode.gflags[sweep & 1] = 0;
        if (prof) sh.tm[kProfRound + 1] += wall_clock64() - tend;
        // bit 1: a rotation above stop_ratio (every rotation when stop_ratio = 0)
        if (!(f & 2)) working = false;
        if (!(gf & 2)) break;
        const bool run = mode.gflags
                             ? __hip_atomic_load(&mode.gflags[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0
                             : want;
        if (run) {
            const unsigned long long tg = prof ? wall_clock64() : 0;
            const bool done =
                gram_converged<Sync, XL>(Wg, ld, m, c, floor2, sv_out, flags + 6, me, nworkers, sync, want);
            if (mode.grid_leader && tid == 0 && mode.gflags) mode.gflags[2] = 0;
            if (prof) sh.tm[kProfRound + 1] += wall_clock64() - tg;
            if (want && done) {
                working = false;
                if (!mode.gflags) break;
            }
        }
    }
    if (prof) {
        sh.tm[2] += wall_clock64() - tp;
        sh.tm[11] += (unsigned long long)sweeps_done;
    }
    if (lane == 0 && nvis) {
        atomicAdd(&sh.vis, nvis * (unsigned long long)m);  // LDS
        atomicAdd(&sh.rots, nrot * (unsigned long long)m);
    }
    // column norms -> sv_out (scratch), then the leader sorts and truncates
    if (mode.active)
        for (int col = me * nwaves + wave; col < c; col += nworkers * nwaves) {
            const double2* x = Wg + (std::size_t)col * ld;
            double s = 0.0;
            for (int row = lane; row < m; row += 64) s += cnorm(gld<XL>(x + row));
            s = wave_sum(s);
            if (lane == 0) sv_out[col] = sqrt(s);
        }
    sync();
    unsigned long long* slot = stats + 4 * (std::size_t)blockIdx.x;
    if (tid == 0) {
        slot[2] += sh.vis;
        slot[3] += sh.rots;
    }
    if (me != 0 || !mode.active) return;
    for (int q = tid; q < c; q += blockDim.x) sh.nrm[q] = gld<XL>(sv_out + q);
    __syncthreads();
    for (int q = tid; q < c; q += blockDim.x) {
        const double v = rank_key(sh.nrm[q]);
        int rank = 0;
        for (int o = 0; o < c; ++o) {
            const double u = rank_key(sh.nrm[o]);
            rank += (u > v) || (u == v && o < q);
        }
        sh.perm[rank] = q;
    }
    __syncthreads();
    if (tid == 0) {
        double total = 0.0;
        for (int q = 0; q < c; ++q) {
            const double s = sh.nrm[sh.perm[q]];
            total += s * s;
        }
        int keep = min(chi, c);
        while (keep > 1 && sh.nrm[sh.perm[keep - 1]] < 1e-14) --keep;
        double disc = 0.0;
        for (int q = keep; q < c; ++q) {
            const double s = sh.nrm[sh.perm[q]];
            disc += s * s;
        }
        if (total > 0) trunc[blockIdx.x] = fmax(trunc[blockIdx.x], disc / total);
        dims[p + 1] = keep;
        slot[0] += 1;
        slot[1] += (unsigned long long)sweeps_done;
    }
    for (int q = tid; q < c; q += blockDim.x) {
        sv_out[q] = sh.nrm[sh.perm[q]];
        perm_out[q] = sh.perm[q];
    }
}

// ---- small SVD held in LDS (one workgroup per update) ----------------------
// For theta with c <= kSmall columns of m <= 16 * RPL rows (bond dimensions up
// to 32, or larger bonds not yet grown; RPL rows per lane): the whole working
// copy sits in LDS and every 16-lane group rotates one column pair per
// tournament round, so a round costs one workgroup barrier instead of the block
// pipeline of svd_blocks. Rotation, convergence test, noise floor, ordering and
// keep rule are those of pair_rotate and svd_blocks (mps.cpp:77-85 for the keep
// rule).
constexpr int kSmall = 64;
constexpr int kSmallRows = kSmall / 16;  // rows per lane of a 16-lane group, m <= 64
constexpr int kSmallTall = 128;          // rows of the tall form (c <= 64, m <= 128)

__device__ inline double sum16(double v) {  // over the aligned 16-lane group
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 1, 64);
    return v;
}

template <int RPL>
__device__ void svd_small(double2* __restrict__ Wg, int ld, int* dims, int p, int chi, int m, int c,
                          double* __restrict__ sv_out, int* __restrict__ perm_out, double* trunc,
                          unsigned long long* stats, double2* A, SvdShared& sh) {
    const int tid = threadIdx.x, l16 = tid & 15, grp = tid >> 4, ngrp = blockDim.x >> 4;
    // columns enter LDS in order of decreasing norm (fewer sweeps); LDS column q
    // holds working column sh.perm[q] and goes back to that column at the end
    for (int col = grp; col < c; col += ngrp) {
        double s2 = 0.0;
        for (int row = l16; row < m; row += 16) s2 += cnorm(Wg[(std::size_t)col * ld + row]);
        s2 = sum16(s2);
        if (l16 == 0) sh.nrm[col] = s2;
    }
    if (tid == 0) {
        sh.vis = 0;
        sh.rots = 0;
    }
    __syncthreads();
    double fro = 0.0;
    for (int q = tid; q < c; q += blockDim.x) {
        const double v = rank_key(sh.nrm[q]);
        fro += v;
        int rank = 0;
        for (int o = 0; o < c; ++o) {
            const double u = rank_key(sh.nrm[o]);
            rank += (u > v) || (u == v && o < q);
        }
        sh.perm[rank] = q;
    }
    fro = wave_sum(fro);
    if ((tid & 63) == 0) sh.red[tid >> 6][0].x = fro;
    __syncthreads();
    for (int idx = tid; idx < m * c; idx += blockDim.x) {
        const int q = idx / m, row = idx % m;
        A[idx] = Wg[(std::size_t)sh.perm[q] * ld + row];
    }
    fro = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) fro += sh.red[w][0].x;
    __syncthreads();
    const double floor2 = noise_floor2(fro), tol = jacobi_tol(m);
    const int cp = c + (c & 1);
    int sweeps = 0;
    unsigned long long nvis = 0, nrot = 0;
    for (int sweep = 0; sweep < 60 && cp >= 2; ++sweep) {
        ++sweeps;
        if (tid == 0) sh.rot = 0;
        // column norms are computed exactly once per sweep and carried across
        // its rotations (|x'|^2 = |x|^2 - t|g|, |y'|^2 = |y|^2 + t|g|), each
        // recomputed once it has dropped 64x below its last exact value (the
        // rounding of the updates stays below ~64 * 63 eps relative), so a
        // visit reduces only the two parts of the pair's inner product
        for (int col = grp; col < c; col += ngrp) {
            const double2* pc = A + (std::size_t)col * m;
            double s2 = 0.0;
            for (int row = l16; row < m; row += 16) s2 += cnorm(pc[row]);
            s2 = sum16(s2);
            if (l16 == 0) {
                sh.cn[col] = s2;
                sh.cr[col] = s2;
            }
        }
        __syncthreads();
        for (int r = 0; r < cp - 1; ++r) {
            for (int k = grp; k < cp / 2; k += ngrp) {
                int i, j;
                rr_pair(r, k, cp, i, j);
                if (i >= c || j >= c) continue;
                double2* px = A + (std::size_t)i * m;
                double2* py = A + (std::size_t)j * m;
                double2 x[RPL], y[RPL];
                double al = sh.cn[i], be = sh.cn[j], gr = 0.0, gi = 0.0;
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int row = l16 + 16 * q;
                    x[q] = row < m ? px[row] : make_double2(0, 0);
                    y[q] = row < m ? py[row] : make_double2(0, 0);
                    gr += x[q].x * y[q].x + x[q].y * y[q].y;  // conj(x) y
                    gi += x[q].x * y[q].y - x[q].y * y[q].x;
                }
                gr = sum16(gr);
                gi = sum16(gi);
                ++nvis;
                if (al <= floor2 || be <= floor2) continue;
                const double gabs = sqrt(gr * gr + gi * gi);
                if (!(gabs > 0.0) || gabs <= tol * sqrt(al * be)) continue;
                const double ig = 1.0 / gabs;
                const double2 ph = make_double2(gr * ig, -gi * ig);
                const double zeta = (be - al) * (0.5 * ig);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                double nx = 0.0, ny = 0.0;
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int row = l16 + 16 * q;
                    const double2 u = x[q], v = cmul(y[q], ph);
                    x[q] = make_double2(cs * u.x - sn * v.x, cs * u.y - sn * v.y);
                    y[q] = make_double2(sn * u.x + cs * v.x, sn * u.y + cs * v.y);
                    nx += cnorm(x[q]);
                    ny += cnorm(y[q]);
                    if (row < m) {
                        px[row] = x[q];
                        py[row] = y[q];
                    }
                }
                double al1 = al - t * gabs, be1 = be + t * gabs;
                // uniform in the group (group-wide values), so the shuffles read active lanes
                const bool rx = al1 < sh.cr[i] * (1.0 / 64), ry = be1 < sh.cr[j] * (1.0 / 64);
                if (rx) al1 = sum16(nx);
                if (ry) be1 = sum16(ny);
                if (l16 == 0) {
                    sh.cn[i] = al1;
                    sh.cn[j] = be1;
                    if (rx) sh.cr[i] = al1;
                    if (ry) sh.cr[j] = be1;
                }
                ++nrot;
                if (l16 == 0) sh.rot = 1;
            }
            __syncthreads();
        }
        const int f = sh.rot;
        __syncthreads();
        if (!f) break;
    }
    // the rotated columns go back to the working copy the reconstruction reads
    // (LDS column q to working column q: the ordering below refers to these)
    for (int idx = tid; idx < m * c; idx += blockDim.x) {
        const int col = idx / m, row = idx % m;
        Wg[(std::size_t)col * ld + row] = A[idx];
    }
    for (int col = grp; col < c; col += ngrp) {
        double s2 = 0.0;
        for (int row = l16; row < m; row += 16) s2 += cnorm(A[(std::size_t)col * m + row]);
        s2 = sum16(s2);
        if (l16 == 0) sh.nrm[col] = sqrt(s2);
    }
    if (l16 == 0 && nvis) {
        atomicAdd(&sh.vis, nvis * (unsigned long long)m);  // LDS
        atomicAdd(&sh.rots, nrot * (unsigned long long)m);
    }
    __syncthreads();
    for (int q = tid; q < c; q += blockDim.x) {  // stable descending rank
        const double v = rank_key(sh.nrm[q]);
        int rank = 0;
        for (int o = 0; o < c; ++o) {
            const double u = rank_key(sh.nrm[o]);
            rank += (u > v) || (u == v && o < q);
        }
        sh.perm[rank] = q;
    }
    __syncthreads();
    unsigned long long* slot = stats + 4 * (std::size_t)blockIdx.x;
    if (tid == 0) {
        double total = 0.0;
        for (int q = 0; q < c; ++q) {
            const double s = sh.nrm[sh.perm[q]];
            total += s * s;
        }
        int keep = min(chi, c);
        while (keep > 1 && sh.nrm[sh.perm[keep - 1]] < 1e-14) --keep;
        double disc = 0.0;
        for (int q = keep; q < c; ++q) {
            const double s = sh.nrm[sh.perm[q]];
            disc += s * s;
        }
        if (total > 0) trunc[blockIdx.x] = fmax(trunc[blockIdx.x], disc / total);
        dims[p + 1] = keep;
        slot[0] += 1;
        slot[1] += (unsigned long long)sweeps;
        slot[2] += sh.vis;
        slot[3] += sh.rots;
    }
    for (int q = tid; q < c; q += blockDim.x) {
        sv_out[q] = sh.nrm[sh.perm[q]];
        perm_out[q] = sh.perm[q];
    }
}

// One update's SVD of a theta with at most kSmall columns and kSmallTall rows
// (the per-update launch sequence): svd_small on one workgroup of kDistThreads
// threads, so every column pair of a tournament round has its own 16-lane group.
__global__ void __launch_bounds__(kDistThreads) k_svd_small(double2* __restrict__ Wg, int ld, int* dims, int p,
                                                            int chi, double* __restrict__ sv_out,
                                                            int* __restrict__ perm_out, double* trunc,
                                                            unsigned long long* stats) {
    extern __shared__ double2 lds[];
    __shared__ SvdShared sh;
    const int R = 2 * dims[p], C = 2 * dims[p + 2];
    const int m = C <= R ? R : C, c = C <= R ? C : R;
    if (m <= kSmall)
        svd_small<kSmallRows>(Wg, ld, dims, p, chi, m, c, sv_out, perm_out, trunc, stats, lds, sh);
    else
        svd_small<kSmallTall / 16>(Wg, ld, dims, p, chi, m, c, sv_out, perm_out, trunc, stats, lds, sh);
}

struct GridSync {
    __device__ void operator()() const { cooperative_groups::this_grid().sync(); }
};

__global__ void __launch_bounds__(kDistThreads) k_svd_dist(double2* __restrict__ Wg, int ld, int* dims, int p,
                                                           int chi, double* __restrict__ sv_out,
                                                           int* __restrict__ perm_out, double* trunc, int* flags,
                                                           unsigned long long* stats, int xcd_stride, double2* aux,
                                                           double2* taus) {
    extern __shared__ double2 lds[];
    __shared__ SvdShared sh;
    GridSync sync;
    // With xcd_stride = 8 only every 8th workgroup works; workgroups are dealt to
    // the XCDs round-robin, so the working ones share one XCD and its L2.
    const bool worker = (blockIdx.x % xcd_stride) == 0;
    const int nworkers = (int)(gridDim.x / xcd_stride);
    const int me = worker ? (int)(blockIdx.x / xcd_stride) : nworkers + (int)blockIdx.x;
    const SvdMode mode;
    svd_blocks(Wg, ld, dims, p, chi, sv_out, perm_out, trunc, flags, stats, me, nworkers, sync, lds, sh, mode, aux,
               taus);
}

// ---- reconstruction (mps.cpp:87-99) ---------------------------------------
// tall (M = theta):        A'[row,k] = W[:,perm k][row]/S_k;  C'[k,col] = sum_row conj(U[row,k]) M[row,col]
// wide (M = theta^dagger): C'[k,col] = conj(W[col, perm k]);  A'[row,k] = sum_col conj(M[col,row]) W[col,perm k] / S_k^2
// With `wts` non-null the right site gets diag(wts) Vh instead of diag(S) Vh
// (used to build benchmark states with prescribed bond weights).
template <bool XL = false>
__device__ void recon_part(double2* sites, const std::size_t* off, const int* cap, const int* dims, int p,
                           const double2* __restrict__ M, const double2* __restrict__ W, int ld,
                           const double* __restrict__ sv, const int* __restrict__ perm, long long gid,
                           long long stride, const double* __restrict__ wts = nullptr) {
    const int Dl = dims[p], keep = gld<XL>(dims + p + 1), Dr = dims[p + 2];
    const int R = 2 * Dl, C = 2 * Dr;
    const bool tall = C <= R;
    const int sA = cap[p + 1], sC = cap[p + 2];
    double2* A = sites + off[p];
    double2* Cn = sites + off[p + 1];
    const long long ndirect = tall ? (long long)R * keep : (long long)keep * C;
    for (long long idx = gid; idx < ndirect; idx += stride) {
        if (tall) {
            const int row = (int)(idx / keep), k = (int)(idx % keep);
            const double s = gld<XL>(sv + k);
            const double2 w = gld<XL>(W + (std::size_t)gld<XL>(perm + k) * ld + row);
            A[(std::size_t)row * sA + k] = s > 0 ? cscale(w, 1.0 / s) : make_double2(0, 0);
        } else {
            const int k = (int)(idx / C), col = (int)(idx % C);
            const int s2 = col / Dr, b = col % Dr;
            const double2 v = cconj(gld<XL>(W + (std::size_t)gld<XL>(perm + k) * ld + col));
            const double s = gld<XL>(sv + k);
            Cn[((std::size_t)k * 2 + s2) * sC + b] = wts ? (s > 0 ? cscale(v, wts[k] / s) : make_double2(0, 0)) : v;
        }
    }
    // products on the matrix cores, one 16 x 16 output tile per wave at a time
    // (gid and stride are thread indices in whole waves):
    //   tall: C'[k, col] = sum_row conj(W[perm k][row]) M[col][row] / S_k  (keep x C, K = R)
    //   wide: A'[row, k] = sum_col conj(M[row][col]) W[perm k][col] / S_k^2 (R x keep, K = C)
    const int lane = (int)(gid & 63), lg = lane >> 4, lc = lane & 15;
    const long long wave = gid >> 6, nwaves = stride >> 6;
    const double2 z = make_double2(0, 0);
    const int rows_out = tall ? keep : R, cols_out = tall ? C : keep, K = tall ? R : C;
    const int ntr = (rows_out + 15) / 16, ntc = (cols_out + 15) / 16;
    for (long long t = wave; t < (long long)ntr * ntc; t += nwaves) {
        const int r0 = (int)(t / ntc) * 16, c0 = (int)(t % ntc) * 16;
        const int ra = r0 + lc, cb = c0 + lc;  // operand row of A, operand column of B for this lane
        // tall: A row ra is singular vector k = ra (column perm k of W), B column cb is M column cb;
        // wide: A row ra is M row ra, B column cb is singular vector k = cb
        const double2* ap;
        const double2* bp;
        bool aok, bok;
        if (tall) {
            aok = ra < keep;
            bok = cb < C;
            ap = W + (std::size_t)(aok ? gld<XL>(perm + ra) : 0) * ld;
            bp = M + (std::size_t)(bok ? cb : 0) * ld;
        } else {
            aok = ra < R;
            bok = cb < keep;
            ap = M + (std::size_t)(aok ? ra : 0) * ld;
            bp = W + (std::size_t)(bok ? gld<XL>(perm + cb) : 0) * ld;
        }
        CTile acc;
        acc.zero();
        for (int k0 = 0; k0 < K; k0 += 16) {
            double2 av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {  // four K steps of loads in flight
                const int kk = k0 + 4 * u + lg;
                const bool ok = kk < K;
                av[u] = (ok && aok) ? cconj(gld<XL>(ap + kk)) : z;
                bv[u] = (ok && bok) ? gld<XL>(bp + kk) : z;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc.mac(av[u], bv[u]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + mfma_out_row(lg, i);
            const double2 d = acc.get(i);
            if (r >= rows_out || cb >= cols_out) continue;
            if (tall) {  // r = k, cb = col
                const double sk = gld<XL>(sv + r);
                const double f = sk > 0 ? (wts ? wts[r] / (sk * sk) : 1.0 / sk) : 0.0;
                const int s2 = cb / Dr, b = cb % Dr;
                Cn[((std::size_t)r * 2 + s2) * sC + b] = cscale(d, f);
            } else {  // r = row, cb = k
                const double sk = gld<XL>(sv + cb);
                A[(std::size_t)r * sA + cb] = sk > 0 ? cscale(d, 1.0 / (sk * sk)) : z;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_recon(double2* sites, const std::size_t* off, const int* cap,
                                               const int* dims, int p, const double2* __restrict__ M,
                                               const double2* __restrict__ W, int ld,
                                               const double* __restrict__ sv, const int* __restrict__ perm,
                                               const double* __restrict__ wts) {
    recon_part(sites, off, cap, dims, p, M, W, ld, sv, perm, (long long)blockIdx.x * blockDim.x + threadIdx.x,
               (long long)gridDim.x * blockDim.x, wts);
}

// ---- environments (mps.cpp:150-178, 218-262) ---------------------------------
// Right: T[a',s,b] = sum_b' R[b,b'] conj(A(a',s,b'));  Rout[a,a'] = sum_{s,b} A(a,s,b) T[a',s,b]
__device__ void envr1_part(const double2* sites, std::size_t o, int st, int Dl, int Dr, const double2* Rin, int ldr,
                           double2* T, long long gid, long long stride) {
    const long long tot = (long long)Dl * 2 * Dr;
    for (long long idx = gid; idx < tot; idx += stride) {
        const int ap = (int)(idx / (2 * Dr)), s = (int)((idx / Dr) % 2), b = (int)(idx % Dr);
        double2 acc = make_double2(0, 0);
        for (int bp = 0; bp < Dr; ++bp)
            acc = cadd(acc, cmul(Rin[(std::size_t)b * ldr + bp], cconj(ld_site(sites, o, st, ap, s, bp))));
        T[idx] = acc;
    }
}
__device__ void envr2_part(const double2* sites, std::size_t o, int st, int Dl, int Dr, const double2* T,
                           double2* Rout, int ldr, long long gid, long long stride) {
    const long long tot = (long long)Dl * Dl;
    for (long long idx = gid; idx < tot; idx += stride) {
        const int a = (int)(idx / Dl), ap = (int)(idx % Dl);
        double2 acc = make_double2(0, 0);
        for (int s = 0; s < 2; ++s)
            for (int b = 0; b < Dr; ++b)
                acc = cadd(acc, cmul(ld_site(sites, o, st, a, s, b), T[((std::size_t)ap * 2 + s) * Dr + b]));
        Rout[(std::size_t)a * ldr + ap] = acc;
    }
}
// Closed-form qubit-0 entropy from site 0 (Dl = 1, Dr) and the right
// environment of site 1 (mps.cpp:263-290).
__device__ double rho0_entropy(const double2* sites, std::size_t o, int st, int Dr, const double2* R1, int ldr) {
    double2 rho[2][2];
    for (int s = 0; s < 2; ++s)
        for (int sp = 0; sp < 2; ++sp) {
            double2 acc = make_double2(0, 0);
            for (int b = 0; b < Dr; ++b)
                for (int bp = 0; bp < Dr; ++bp)
                    acc = cadd(acc, cmul(cmul(ld_site(sites, o, st, 0, s, b), cconj(ld_site(sites, o, st, 0, sp, bp))),
                                         R1[(std::size_t)b * ldr + bp]));
            rho[s][sp] = acc;
        }
    const double tr = rho[0][0].x + rho[1][1].x;
    double ent = 0.0;
    if (tr > 0) {
        const double a = rho[0][0].x / tr, d = rho[1][1].x / tr;
        const double2 of = make_double2(rho[0][1].x / tr, rho[0][1].y / tr);
        const double off2 = of.x * of.x + of.y * of.y;
        const double mean = (a + d) / 2.0;
        const double rad = sqrt(fmax(0.0, ((a - d) / 2) * ((a - d) / 2) + off2));
        const double l1 = fmin(1.0, fmax(0.0, mean + rad)), l2 = fmin(1.0, fmax(0.0, mean - rad));
        if (l1 > 1e-12) ent -= l1 * log2(l1);
        if (l2 > 1e-12) ent -= l2 * log2(l2);
    }
    return ent;
}

__global__ void __launch_bounds__(256) k_envr1(const double2* sites, const std::size_t* off, const int* cap,
                                               const int* dims, int j, const double2* Rin, int ldr, double2* T) {
    const int Dl = dims[j], Dr = dims[j + 1], st = cap[j + 1];
    const std::size_t o = off[j];
    const long long tot = (long long)Dl * 2 * Dr;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
         idx += (long long)gridDim.x * blockDim.x) {
        const int ap = (int)(idx / (2 * Dr)), s = (int)((idx / Dr) % 2), b = (int)(idx % Dr);
        double2 acc = make_double2(0, 0);
        for (int bp = 0; bp < Dr; ++bp)
            acc = cadd(acc, cmul(Rin[(std::size_t)b * ldr + bp], cconj(ld_site(sites, o, st, ap, s, bp))));
        T[idx] = acc;
    }
}
__global__ void __launch_bounds__(256) k_envr2(const double2* sites, const std::size_t* off, const int* cap,
                                               const int* dims, int j, const double2* T, double2* Rout, int ldr) {
    const int Dl = dims[j], Dr = dims[j + 1], st = cap[j + 1];
    const std::size_t o = off[j];
    const long long tot = (long long)Dl * Dl;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
         idx += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(idx / Dl), ap = (int)(idx % Dl);
        double2 acc = make_double2(0, 0);
        for (int s = 0; s < 2; ++s)
            for (int b = 0; b < Dr; ++b)
                acc = cadd(acc, cmul(ld_site(sites, o, st, a, s, b), T[((std::size_t)ap * 2 + s) * Dr + b]));
        Rout[(std::size_t)a * ldr + ap] = acc;
    }
}
// Left: T[a,s,b'] = sum_a' L[a,a'] conj(A(a',s,b'));  Lout[b,b'] = sum_{s,a} zf(s) A(a,s,b) T[a,s,b']
__global__ void __launch_bounds__(256) k_envl1(const double2* sites, const std::size_t* off, const int* cap,
                                               const int* dims, int j, const double2* Lin, int ldl, double2* T) {
    const int Dl = dims[j], Dr = dims[j + 1], st = cap[j + 1];
    const std::size_t o = off[j];
    const long long tot = (long long)Dl * 2 * Dr;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
         idx += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(idx / (2 * Dr)), s = (int)((idx / Dr) % 2), bp = (int)(idx % Dr);
        double2 acc = make_double2(0, 0);
        for (int ap = 0; ap < Dl; ++ap)
            acc = cadd(acc, cmul(Lin[(std::size_t)a * ldl + ap], cconj(ld_site(sites, o, st, ap, s, bp))));
        T[idx] = acc;
    }
}
__global__ void __launch_bounds__(256) k_envl2(const double2* sites, const std::size_t* off, const int* cap,
                                               const int* dims, int j, int withz, const double2* T, double2* Lout,
                                               int ldl) {
    const int Dl = dims[j], Dr = dims[j + 1], st = cap[j + 1];
    const std::size_t o = off[j];
    const long long tot = (long long)Dr * Dr;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
         idx += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(idx / Dr), bp = (int)(idx % Dr);
        double2 acc = make_double2(0, 0);
        for (int s = 0; s < 2; ++s) {
            double2 part = make_double2(0, 0);
            for (int a = 0; a < Dl; ++a)
                part = cadd(part, cmul(ld_site(sites, o, st, a, s, b), T[((std::size_t)a * 2 + s) * Dr + bp]));
            acc = (withz && s == 1) ? csub(acc, part) : cadd(acc, part);
        }
        Lout[(std::size_t)b * ldl + bp] = acc;
    }
}

__global__ void k_set_unit(double2* e) { e[0] = make_double2(1.0, 0.0); }

// Qubit-0 reduced density matrix from the right environment of site 1 and the
// closed-form 2x2 entropy (mps.cpp:263-290).
__global__ void k_rho0(const double2* sites, const std::size_t* off, const int* cap, const int* dims,
                       const double2* R1, int ldr, double* out, int slot, const int* par = nullptr) {
    if (threadIdx.x != 0) return;
    if (par) slot += par[1] + par[0];  // replayed step: its recorded slot
    const int Dr = dims[1], st = cap[1];
    double2 rho[2][2];
    for (int s = 0; s < 2; ++s)
        for (int sp = 0; sp < 2; ++sp) {
            double2 acc = make_double2(0, 0);
            for (int b = 0; b < Dr; ++b)
                for (int bp = 0; bp < Dr; ++bp)
                    acc = cadd(acc, cmul(cmul(ld_site(sites, off[0], st, 0, s, b),
                                              cconj(ld_site(sites, off[0], st, 0, sp, bp))),
                                         R1[(std::size_t)b * ldr + bp]));
            rho[s][sp] = acc;
        }
    const double tr = rho[0][0].x + rho[1][1].x;
    double ent = 0.0;
    if (tr > 0) {
        const double a = rho[0][0].x / tr, d = rho[1][1].x / tr;
        const double2 o = make_double2(rho[0][1].x / tr, rho[0][1].y / tr);
        const double off2 = o.x * o.x + o.y * o.y;
        const double mean = (a + d) / 2.0;
        const double rad = sqrt(fmax(0.0, ((a - d) / 2) * ((a - d) / 2) + off2));
        const double l1 = fmin(1.0, fmax(0.0, mean + rad)), l2 = fmin(1.0, fmax(0.0, mean - rad));
        if (l1 > 1e-12) ent -= l1 * log2(l1);
        if (l2 > 1e-12) ent -= l2 * log2(l2);
    }
    out[slot] = ent;
}

// sum_{b,b'} X[b,b'] R[b,b'] over a d x d block -> out[slot]
__global__ void __launch_bounds__(256) k_close(const double2* X, const double2* Rm, int ld, const int* dims,
                                               int bond, double2* out, int slot) {
    __shared__ double red[4][2];
    const int d = dims[bond];
    double re = 0, im = 0;
    for (int idx = threadIdx.x; idx < d * d; idx += blockDim.x) {
        const int b = idx / d, bp = idx % d;
        const double2 v = cmul(X[(std::size_t)b * ld + bp], Rm[(std::size_t)b * ld + bp]);
        re += v.x;
        im += v.y;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = re;
        red[threadIdx.x >> 6][1] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a += red[w][0];
            b += red[w][1];
        }
        out[slot] = make_double2(a, b);
    }
}

// <x|psi> for a fixed configuration (mps.cpp:197-213), one workgroup.
__global__ void __launch_bounds__(256) k_amplitude(const double2* sites, const std::size_t* off, const int* cap,
                                                   const int* dims, const int8_t* spins, int n, double2* out) {
    __shared__ double2 v[2][256];
    if (threadIdx.x == 0) v[0][0] = make_double2(1, 0);
    __syncthreads();
    int cur = 0;
    for (int i = 0; i < n; ++i) {
        const int Dl = dims[i], Dr = dims[i + 1], st = cap[i + 1];
        const int s = spins[i] > 0 ? 0 : 1;
        for (int b = threadIdx.x; b < Dr; b += blockDim.x) {
            double2 acc = make_double2(0, 0);
            for (int a = 0; a < Dl; ++a) acc = cadd(acc, cmul(v[cur][a], ld_site(sites, off[i], st, a, s, b)));
            v[cur ^ 1][b] = acc;
        }
        __syncthreads();
        cur ^= 1;
    }
    if (threadIdx.x == 0) *out = v[cur][0];
}

// Random saturated state: Gaussian entries from a counter hash, scaled by
// exp(-b/decay) on the right bond index and 1/sqrt(2 Dr) per site.
__device__ inline double u01(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__global__ void k_randomize(double2* sites, const std::size_t* off, const int* cap, int n, unsigned long long seed,
                            double decay) {
    const int i = blockIdx.y;
    const int Dl = cap[i], Dr = cap[i + 1];
    const double norm = 1.0 / sqrt(2.0 * Dr);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < Dl * 2 * Dr; idx += gridDim.x * blockDim.x) {
        const int b = idx % Dr;
        const unsigned long long key = seed * 0x100000001B3ull + (unsigned long long)off[i] + idx;
        const double u1 = u01(2 * key), u2 = u01(2 * key + 1);
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(6.283185307179586 * u2, &sn, &cs);
        const double w = norm * exp(-b / decay);
        sites[off[i] + idx] = make_double2(rad * cs * w, rad * sn * w);
    }
}

// theta of update p (theta_mfma) by the waves of `nwg` workgroups, this one
// being number `rank`.
__device__ void theta_tiles(const double2* __restrict__ sites, const std::size_t* off, const int* cap,
                            const int* dims, int p, const Gate4& g, double2* __restrict__ M, double2* __restrict__ W,
                            int ld, int rank, int nwg) {
    const int wpg = (int)(blockDim.x >> 6);
    theta_mfma(sites, off, cap, dims, p, g, M, W, ld, (long long)rank * wpg + (threadIdx.x >> 6),
               (long long)nwg * wpg);
}

// ---------------------------------------------------------------------------
// Lockstep executor. One cooperative launch holds G groups of W workgroups
// (group g = blockIdx % G; with G a multiple of 8 a group sits on one XCD, as
// workgroups are dealt to the 8 XCDs round-robin). The host schedules the
// two-site updates of several Trotter steps into super-steps of at most G
// mutually independent updates (an update runs only after every earlier update
// on either of its sites), so consecutive steps overlap. Each super-step is
//   1. site snapshots due at this super-step (all workgroups), theta of each
//      group's update (its workgroups);
//   2. block-Jacobi SVDs of all groups in lockstep (one grid barrier per
//      tournament round, sweeps until no group rotates);
//   3. the group leaders sort and truncate;
//   4. reconstruction of each group's two sites.
// Only the cooperative-groups grid barrier synchronises workgroups. A snapshot
// copies a site after the last update of its step and before the next update
// touching it changes the site (sites change only in phase 4, bond dims at the
// end of phase 3), with the single-site gates still pending applied.
struct ExecUpdate {
    int site;     // -1: idle slot
    int inst;     // instance (MPS) the update belongs to
    int pred[2];  // latest earlier updates on sites site, site + 1 (flat indices; -1: none)
    Gate4 g;      // two-site gate with pending single-site gates folded in
};
// One MPS advanced by the executor (several in a batched run).
struct ExecInst {
    double2* sites;
    const std::size_t* off;
    const int* cap;
    int* dims;
    int chi;
    double* trunc;               // per-workgroup slots
    unsigned long long* stats;   // per-workgroup slots
};
struct ExecSnap {
    int site;
    int slot;       // snapshot index (recorded step)
    double2 g[4];   // 2x2 gate applied to the copy
};
struct ExecArgs {
    const ExecInst* insts;     // instance 0 also owns the snapshots
    const ExecUpdate* ups;     // updates of super-step st: ups[up_begin[st] .. up_begin[st+1])
    const int* up_begin;       // [nsteps + 1]
    const int* rounds;         // [nsteps] tournament rounds of the widest SVD
    const int* snap_begin;     // [nsteps + 1]
    const ExecSnap* snaps;
    int nsteps;
    int G;
    int n;              // sites of instance 0 (snapshots)
    double2* mat;
    double2* work;
    int ld;
    double* sv;
    int* perm;
    int* svd_flags;     // [G][8]: rotation flags [0, 1], rotation counts [4, 5], check flag [6]
    int* gflags;        // [3]: grid-wide rotation flags (per sweep parity), a convergence check asked for
    double2* snap;
    std::size_t snap_stride;
    int* snap_dims;     // [slots][n + 1]
    double2* aux;       // [G][ld * ld] LQ reflectors
    double2* taus;      // [G][ld]
    int small;          // solo layout: thetas within kSmall x kSmall take svd_small (dynamic LDS fits it)
    int* group_bar;     // [G][2] arrival count, generation
    int* err;           // raised by a GroupSync wait that gave up
    long long spin_limit;  // GroupSync polls before giving up
    int naps;              // 64-cycle sleeps between the polls of a wait
    double stop_ratio;     // SvdMode::stop_ratio of every SVD
    int rows;              // SvdMode::rows of every SVD
    int gram;              // SvdMode::gram of the non-lockstep SVDs
    int step_prof;         // SvdMode::step_prof of the profiled workgroup
    int prof_on;           // workgroup 0 records the phase profile
    int* epochs;        // [G][W][64] round counters of the point-to-point rounds (per-group layout)
    int dataflow;       // per-group layout: no super-step barriers; an update waits for the
                        // done flags of its predecessors instead
    int* done;          // [updates] done flags of the dataflow order
    unsigned long long* done_clock;  // [updates + 1] or null (dataflow): wall clock at each update's
                                     // done flag, then at the start
    // dataflow order: snaps[0, npro) are taken before any update (their sites have
    // none earlier in the launch); snaps[usnap_begin[f], usnap_begin[f + 1]) by
    // update f after its reconstruction (it is the latest update on their site)
    const int* usnap_begin;  // [updates + 1]
    int npro;
    // dataflow order with a shared queue: queue[0] is the next claim, queue[1 + g]
    // group g's latest claim; claim k is update qorder[k], an order in which every
    // update comes after its predecessors, so every claimed update's predecessors
    // were claimed earlier by running groups.
    // Null: group g runs slot g of every super-step.
    int* queue;
    const int* qorder;  // [updates]: the flat index of the update handed out as claim k
    int nup;  // updates of the launch
    unsigned long long* prof;  // [kProfSlots] phase profile of workgroup 0 (accumulated)
    unsigned long long* ss_clock;  // [nsteps + 1] or null: wall clock at the end of each super-step, then at the start
};

// Site snapshot of instance 0 for the entropy trace: site i with the pending
// single-site gate applied, into its slot (threads idx0, idx0 + stride, ...).
__device__ void take_snapshot(const ExecSnap& sn, const ExecInst& I0, double2* snap, std::size_t snap_stride,
                              int* snap_dims, int n, long long idx0, long long stride) {
    const int i = sn.site;
    const int Dl = I0.dims[i], Dr = I0.dims[i + 1], stc = I0.cap[i + 1];
    const double2* src = I0.sites + I0.off[i];
    double2* dst = snap + (std::size_t)sn.slot * snap_stride + I0.off[i];
    for (long long idx = idx0; idx < (long long)Dl * Dr; idx += stride) {
        const int x = (int)(idx / Dr), b = (int)(idx % Dr);
        const double2 v0 = src[((std::size_t)x * 2 + 0) * stc + b];
        const double2 v1 = src[((std::size_t)x * 2 + 1) * stc + b];
        dst[((std::size_t)x * 2 + 0) * stc + b] = cadd(cmul(sn.g[0], v0), cmul(sn.g[1], v1));
        dst[((std::size_t)x * 2 + 1) * stc + b] = cadd(cmul(sn.g[2], v0), cmul(sn.g[3], v1));
    }
    if (idx0 == 0) {
        int* sd = snap_dims + (std::size_t)sn.slot * (n + 1);
        sd[i] = Dl;
        sd[i + 1] = Dr;
    }
}

// Barrier of the nwg co-resident workgroups of one executor group (the
// cooperative launch keeps every workgroup resident). bar[0] counts arrivals,
// bar[1] is the generation: the last arriver resets the count and advances the
// generation, the others poll it. Every wait ends: when the generation moves,
// when another workgroup has raised *err, or after kSpinLimit polls, which
// raise *err (the host reports it) instead of hanging the device.
constexpr long long kSpinLimit = 1ll << 24;
constexpr long long kDoneWaitFactor = 64;
template <bool XL = false>
struct GroupSync {
    int* bar;
    int nwg;
    int* err;
    long long limit;  // polls before giving up (kSpinLimit unless a test lowers it)
    int naps = 1;     // 64-cycle sleeps between polls
    __device__ void operator()() const {
        __builtin_amdgcn_s_waitcnt(0);  // every wave's stores complete in the L2 before the arrival
        __syncthreads();
        if (threadIdx.x == 0) {
            if constexpr (XL) {
                // one XCD: bar[0] counts arrivals monotonically (zeroed per launch);
                // the barrier of generation k completes at (k + 1) * nwg arrivals.
                // Relaxed device-scope atomics and polls are served by the shared L2;
                // the data behind the barrier is read with L1-bypassing loads.
                const int prev = __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int target = (prev / nwg + 1) * nwg;
                for (long long spins = 0;
                     __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++spins) {
                    if (spins >= limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            } else {
                // thread 0's acq_rel arrival releases every wave's stores (complete in
                // the L2 after the s_waitcnt and barrier above) with the L2 at agent
                // scope; the last arrival also acquires, the others acquire with
                // their polls
                const int gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int prev = __hip_atomic_fetch_add(bar, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                if (prev == nwg - 1) {
                    __hip_atomic_store(bar, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(bar + 1, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    // relaxed polls (an acquire poll would invalidate the XCD's L2 each
                    // time), then one acquire fence that synchronises with the last
                    // arrival's release; the error word is read every 32nd poll
                    bool gave_up = false;
                    for (long long spins = 0;
                         __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen; ++spins) {
                        if (spins >= limit ||
                            ((spins & 31) == 31 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            gave_up = true;
                            break;
                        }
                        for (int q = 0; q < 2 * naps; ++q) __builtin_amdgcn_s_sleep(1);
                    }
                    if (!gave_up) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
            }
        }
        __syncthreads();
    }
};

// Barrier of one workgroup for code written against a barrier functor.
struct BlockSync {
    __device__ void operator()() const {
        // __syncthreads carries workgroup-scope release and acquire fences, which
        // order global memory among the workgroup's waves (they share the CU's
        // L1); other workgroups read this data only after a grid barrier
        __syncthreads();
    }
};

// Layouts, one instantiation each so that every one stays within its registers:
// kExecLockstep (grouped, all groups' SVDs advance in lockstep on the grid
// barrier), kExecGroupSync (grouped, each group's update on the group's own
// barrier, GroupSync), kExecSolo (one workgroup per update).
constexpr int kExecLockstep = 0, kExecGroupSync = 1, kExecSolo = 2;
// Rows of a Jacobi round per thread in the grouped layouts (workgroups of 512 /
// kExecRows threads; the solo layout always runs 512 threads with one row each).
// The row layout, LQ and convergence check are written for 1 or 2 rows per
// thread, but 2 is switched off: its first run (n = 12, chi = 64, eight workers
// per group) ended in an illegal memory access whose cause has not been found.
constexpr int kExecRows = 1;
__host__ __device__ constexpr int exec_threads(int layout) {
    return layout == kExecSolo ? kDistThreads : kDistThreads / kExecRows;
}
// XL (grouped layout only): the groups sit one per XCD and synchronise
// XCD-locally (GroupSync<true>, L1-bypassing loads of the group's data).
template <int LAYOUT, bool XL = false>
__global__ void __launch_bounds__(exec_threads(LAYOUT)) k_tebd_lockstep(ExecArgs a) {
    extern __shared__ double2 lds[];
    __shared__ SvdShared sh;
    __shared__ int s_claim;
    const int tid = threadIdx.x;
    const int G = a.G;
    const int grp = (int)(blockIdx.x % G), rank = (int)(blockIdx.x / G), nwg = (int)(gridDim.x / G);
    GridSync sync;
    const std::size_t mat_sz = (std::size_t)a.ld * a.ld;
    double2* M = a.mat + grp * mat_sz;
    double2* W = a.work + grp * mat_sz;
    double* sv = a.sv + (std::size_t)grp * a.ld;
    int* perm = a.perm + (std::size_t)grp * a.ld;
    int* sflags = a.svd_flags + 8 * grp;
    const long long gid = (long long)rank * blockDim.x + tid, gstride = (long long)nwg * blockDim.x;
    const long long tid_all = (long long)blockIdx.x * blockDim.x + tid;
    const long long all = (long long)gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && tid == 0) {
        a.gflags[0] = 0;
        a.gflags[1] = 0;
        a.gflags[2] = 0;
    }
    if constexpr (XL) {
        // every group must sit on one XCD (blockIdx mod 8 = XCC id, G a multiple of
        // 8); otherwise raise bit 2 of the error word and do nothing
        if (tid == 0 && xcc_id() != (blockIdx.x & 7u))
            __hip_atomic_fetch_or(a.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sync();
        if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2) return;
    }
    // phase profile of workgroup 0 (wall-clock ticks): 0 snapshots + theta,
    // 1 LQ preconditioning, 2 Jacobi sweeps, 3 whole SVD, 4 wait after the SVD,
    // 5 reconstruction, 6 wait at the super-step end (dataflow order: claims and
    // waits for predecessors), 7 updates run by the group, 8 waits at
    // tournament-round barriers, 9 rounds, 10 kernel, 11 sweeps
    const bool prof = a.prof_on && blockIdx.x == 0 && tid == 0;
    if (prof)
        for (int q = 0; q < kProfSlots; ++q) sh.tm[q] = 0;
    const unsigned long long t_start = prof ? wall_clock64() : 0;
    if (blockIdx.x == 0 && tid == 0) {  // the launch's start for the completion times
        if (a.ss_clock) a.ss_clock[a.nsteps] = wall_clock64();
        if (a.done_clock) a.done_clock[a.up_begin[a.nsteps]] = wall_clock64();
    }
    unsigned long long tp = t_start;
    auto mark = [&](int slot) {
        if (prof) {
            const unsigned long long t = wall_clock64();
            sh.tm[slot] += t - tp;
            tp = t;
        }
    };
    const ExecInst& I0 = a.insts[0];
    if (a.dataflow && a.npro > 0) {  // snapshots of sites no update of the launch touches first
        for (int q = 0; q < a.npro; ++q) take_snapshot(a.snaps[q], I0, a.snap, a.snap_stride, a.snap_dims, a.n, tid_all, all);
        sync();
    }
    const bool queue = LAYOUT == kExecGroupSync && a.queue != nullptr;
    for (int st = 0; queue || st < a.nsteps; ++st) {
        int claim = 0;
        if constexpr (LAYOUT == kExecGroupSync) {
            if (queue) {  // the group's next update from the shared queue
                if (rank == 0 && tid == 0)
                    __hip_atomic_store(&a.queue[1 + grp],
                                       __hip_atomic_fetch_add(&a.queue[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                GroupSync<XL>{a.group_bar + 2 * grp, nwg, a.err, a.spin_limit, a.naps}();
                if (tid == 0) s_claim = __hip_atomic_load(&a.queue[1 + grp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                claim = s_claim;
                // rank 0 writes the next claim only after the update's group barriers,
                // which every workgroup reaches after reading this one
                if (claim >= a.nup) break;
            }
        }
        const bool active = queue || grp < a.up_begin[st + 1] - a.up_begin[st];
        const int fl = queue ? a.qorder[claim] : (active ? a.up_begin[st] + grp : 0);  // flat update index
        const ExecUpdate& u = a.ups[fl];
        const ExecInst& I = a.insts[active ? u.inst : 0];
        // 1. snapshots (whole grid, instance 0; the dataflow order takes them with
        // the updates), theta (group)
        if (!a.dataflow)
            for (int q = a.snap_begin[st]; q < a.snap_begin[st + 1]; ++q)
                take_snapshot(a.snaps[q], I0, a.snap, a.snap_stride, a.snap_dims, a.n, tid_all, all);
        if constexpr (LAYOUT == kExecSolo) {
            // the snapshots read sites that this super-step's reconstructions may rewrite
            if (a.snap_begin[st + 1] > a.snap_begin[st]) sync();
            if (active) {
                BlockSync bs;
                theta_tiles(I.sites, I.off, I.cap, I.dims, u.site, u.g, M, W, a.ld, 0, 1);
                bs();
                const int R = 2 * I.dims[u.site], C = 2 * I.dims[u.site + 2];
                const int m = C <= R ? R : C, c = C <= R ? C : R;
                // thetas within the kSmall x kSmall LDS area: 64 x 64, or up to 128 x 32
                if (a.small && m <= kSmall && c <= kSmall) {
                    svd_small<kSmallRows>(W, a.ld, I.dims, u.site, I.chi, m, c, sv, perm, I.trunc, I.stats, lds, sh);
                } else if (a.small && m <= kSmallTall && c <= kSmall && m * c <= kSmall * kSmall) {
                    svd_small<kSmallTall / 16>(W, a.ld, I.dims, u.site, I.chi, m, c, sv, perm, I.trunc, I.stats, lds,
                                               sh);
                } else {
                    SvdMode mode;  // own tournament length, own convergence
                    mode.stop_ratio = a.stop_ratio;
                    mode.rows = a.rows != 0;
                    mode.gram = a.gram != 0;
                    svd_blocks(W, a.ld, I.dims, u.site, I.chi, sv, perm, I.trunc, sflags, I.stats, 0, 1, bs, lds,
                               sh, mode, (a.aux ? a.aux + grp * mat_sz : nullptr), a.taus + (std::size_t)grp * a.ld);
                }
                bs();
                recon_part(I.sites, I.off, I.cap, I.dims, u.site, M, W, a.ld, sv, perm, tid, blockDim.x);
            }
            sync();  // the super-step's sites are final
        } else if constexpr (LAYOUT == kExecGroupSync) {
            // the snapshots read sites that this super-step's reconstructions may rewrite
            if (!a.dataflow && a.snap_begin[st + 1] > a.snap_begin[st]) sync();
            if (active) {
                if (a.dataflow) {
                    // the sites of this update are final once the latest earlier
                    // updates on them are done (each of those waited for its own)
                    if (tid == 0) {
                        for (int q = 0; q < 2; ++q)
                            // a group with no update in the super-steps before this one
                            // arrives here early and may wait as long as the critical path
                            // ahead of it (seconds at n = 128, chi = 256), so this wait
                            // allows kDoneWaitFactor times the polls of a barrier wait
                            if (u.pred[q] >= 0)
                                wait_counter(&a.done[u.pred[q]], 1, a.err, a.spin_limit * kDoneWaitFactor);
                        // each wait ends with an acquire fence: the predecessors' sites are visible
                    }
                    __syncthreads();
                    mark(6);
                }
                // theta, SVD and reconstruction of this group's update on the
                // group's own barrier: own tournament length and convergence
                GroupSync<XL> gs{a.group_bar + 2 * grp, nwg, a.err, a.spin_limit, a.naps};
                theta_tiles(I.sites, I.off, I.cap, I.dims, u.site, u.g, M, W, a.ld, rank, nwg);
                mark(0);
                SvdMode mode;
                mode.prof = a.prof_on && blockIdx.x == 0;
                mode.stop_ratio = a.stop_ratio;
                mode.rows = a.rows != 0;
                mode.gram = a.gram != 0;
                mode.step_prof = a.step_prof != 0;
                mode.epochs = a.epochs ? a.epochs + (std::size_t)grp * nwg * 64 : nullptr;
                mode.err = a.err;
                mode.spin_limit = a.spin_limit;
                mode.naps = a.naps;
                svd_blocks<GroupSync<XL>, XL, kExecRows>(W, a.ld, I.dims, u.site, I.chi, sv, perm, I.trunc, sflags, I.stats, rank,
                                              nwg, gs, lds, sh, mode, (a.aux ? a.aux + grp * mat_sz : nullptr),
                                              a.taus + (std::size_t)grp * a.ld);
                mark(3);
                gs();
                mark(4);
                recon_part<XL>(I.sites, I.off, I.cap, I.dims, u.site, M, W, a.ld, sv, perm, gid, gstride);
                mark(5);
                if (prof) sh.tm[7] += 1;
                if (a.dataflow) {
                    gs();  // every workgroup's part of the reconstruction is released
                    if (a.usnap_begin[fl + 1] > a.usnap_begin[fl]) {
                        // the snapshots of the sites this update leaves final for a step
                        for (int q = a.usnap_begin[fl]; q < a.usnap_begin[fl + 1]; ++q)
                            take_snapshot(a.snaps[q], I0, a.snap, a.snap_stride, a.snap_dims, a.n, gid, gstride);
                        gs();  // read before a later update on the sites may rewrite them
                    }
                    if (rank == 0 && tid == 0) {
                        __hip_atomic_store(&a.done[fl], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        if (a.done_clock) a.done_clock[fl] = wall_clock64();
                    }
                    mark(4);
                }
            }
            if (!a.dataflow) {
                sync();  // the super-step's sites are final
                mark(6);
            }
        } else {
            if (active) theta_tiles(I.sites, I.off, I.cap, I.dims, u.site, u.g, M, W, a.ld, rank, nwg);
            mark(0);
            // 2-3. lockstep SVD; svd_blocks opens with a grid barrier
            SvdMode mode;
            mode.stop_ratio = a.stop_ratio;
            mode.rows = a.rows != 0;
            mode.step_prof = a.step_prof != 0;
            mode.rounds = a.rounds[st];
            mode.gflags = a.gflags;
            mode.gram = a.gram != 0;
            mode.grid_leader = blockIdx.x == 0;
            mode.active = active;
            mode.lq_panels = (a.ld + 15) / 16;
            mode.prof = a.prof_on && blockIdx.x == 0;
            svd_blocks<GridSync, false, kExecRows>(W, a.ld, I.dims, active ? u.site : 0, I.chi, sv, perm, I.trunc, sflags,
                                                   I.stats, rank, nwg, sync,
                       lds, sh, mode, (a.aux ? a.aux + grp * mat_sz : nullptr), a.taus + (std::size_t)grp * a.ld);
            mark(3);
            sync();
            mark(4);
            // 4. reconstruction
            if (active) recon_part(I.sites, I.off, I.cap, I.dims, u.site, M, W, a.ld, sv, perm, gid, gstride);
            mark(5);
            if (prof && active) sh.tm[7] += 1;
            sync();
            mark(6);
        }
        // every layout has just passed the super-step's closing grid barrier (the
        // dataflow order has none and records no clock)
        if (a.ss_clock && blockIdx.x == 0 && tid == 0) a.ss_clock[st] = wall_clock64();
    }
    if (prof) {
        sh.tm[10] += wall_clock64() - t_start;
        for (int q = 0; q < kProfSlots; ++q) a.prof[q] += sh.tm[q];
    }
}

// Applies one 2x2 gate per site (flushes the executor's pending single-site gates).
__global__ void __launch_bounds__(256) k_apply_site_gates(double2* sites, const std::size_t* off, const int* cap,
                                                          const int* dims, const double2* gates) {
    const int i = blockIdx.x;
    const int Dl = dims[i], Dr = dims[i + 1], st = cap[i + 1];
    const double2 g00 = gates[4 * i], g01 = gates[4 * i + 1], g10 = gates[4 * i + 2], g11 = gates[4 * i + 3];
    double2* t = sites + off[i];
    for (int idx = threadIdx.x; idx < Dl * Dr; idx += blockDim.x) {
        const int x = idx / Dr, b = idx % Dr;
        double2* p0 = t + ((std::size_t)x * 2 + 0) * st + b;
        double2* p1 = t + ((std::size_t)x * 2 + 1) * st + b;
        const double2 v0 = *p0, v1 = *p1;
        *p0 = cadd(cmul(g00, v0), cmul(g01, v1));
        *p1 = cadd(cmul(g10, v0), cmul(g11, v1));
    }
}

int64_t grid_for(long long work, int block) {
    long long g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65535) g = 65535;
    return g;
}

}  // namespace

Tebd::Tebd(const TebdModel& model, int chi, double dt, int zz_substeps)
    : n_(model.n), chi_(chi), ns_(zz_substeps < 1 ? 1 : zz_substeps), dt_(dt), model_(model) {
    if (n_ < 1) throw Error(-1, "TEBD needs n >= 1");
    if (chi_ < 1) throw Error(-1, "TEBD needs chi >= 1");
    if (chi_ > 256) throw Error(-2, "this build supports chi <= 256");
    for (std::size_t e = 0; e < model_.ew.size(); ++e)
        if (model_.ei[e] < 0 || model_.ej[e] < 0 || model_.ei[e] >= n_ || model_.ej[e] >= n_)
            throw Error(-1, "edge index out of range");
    cap_.resize(n_ + 1);
    for (int b = 0; b <= n_; ++b) {
        const int e = std::min(b, n_ - b);
        cap_[b] = e >= 30 ? chi_ : std::min<long long>(chi_, 1ll << e);
    }
    off_.resize(n_ + 1);
    std::size_t tot = 0;
    for (int i = 0; i < n_; ++i) {
        off_[i] = tot;
        tot += (std::size_t)cap_[i] * 2 * cap_[i + 1];
    }
    off_[n_] = tot;
    TQ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    sites_.alloc(tot);
    dims_.alloc(n_ + 1);
    int maxcap = 1;
    for (int c : cap_) maxcap = std::max(maxcap, c);
    maxcap_ = maxcap;
    const int ld = 2 * maxcap;
    ld_ = ld;
    flags_.alloc(8);  // rotation flags [0, 1], rotation counts [4, 5], check flag [6]
    stats_.alloc(4 * (std::size_t)kStatSlots);
    TQ_HIP(hipMemset(stats_.p, 0, 4 * (std::size_t)kStatSlots * sizeof(unsigned long long)));
    mat_.alloc((std::size_t)ld * ld);
    work_.alloc((std::size_t)ld * ld);
    aux_.alloc((std::size_t)ld * ld);
    taus_.alloc(ld);
    sv_.alloc(ld);
    perm_.alloc(ld);
    trunc_.alloc(kStatSlots);
    h_.alloc(n_);
    TQ_HIP(hipMemcpy(h_.p, model_.h.data(), n_ * sizeof(double), hipMemcpyHostToDevice));
    if (!model_.bias.empty()) {
        bias_.alloc(n_);
        TQ_HIP(hipMemcpy(bias_.p, model_.bias.data(), n_ * sizeof(double), hipMemcpyHostToDevice));
    }
    envA_.alloc((std::size_t)maxcap * maxcap);
    envB_.alloc((std::size_t)maxcap * maxcap);
    envT_.alloc((std::size_t)maxcap * 2 * maxcap);
    // device copies of the layout tables live right after the dims array
    caps_dev_.alloc(n_ + 1);
    off_dev_.alloc(n_ + 1);
    TQ_HIP(hipMemcpy(caps_dev_.p, cap_.data(), (n_ + 1) * sizeof(int), hipMemcpyHostToDevice));
    TQ_HIP(hipMemcpy(off_dev_.p, off_.data(), (n_ + 1) * sizeof(std::size_t), hipMemcpyHostToDevice));
    TQ_HIP(hipFuncSetAttribute((const void*)k_svd, hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    TQ_HIP(hipFuncSetAttribute((const void*)k_svd_dist, hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    for (const void* k : {(const void*)k_tebd_lockstep<kExecLockstep>, (const void*)k_tebd_lockstep<kExecGroupSync>,
                          (const void*)k_tebd_lockstep<kExecGroupSync, true>, (const void*)k_tebd_lockstep<kExecSolo>})
        TQ_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024));
    TQ_HIP(hipFuncSetAttribute((const void*)k_svd_small, hipFuncAttributeMaxDynamicSharedMemorySize,
                               kSmall * kSmallTall * (int)sizeof(double2)));
    const char* fd = std::getenv("TESSERA_FORCE_DIST_SVD");  // test knob: distributed SVD at every size
    force_dist_ = fd && fd[0] == '1';
    const char* ss = std::getenv("TESSERA_SMALL_SVD");  // test knob: 0 keeps small thetas on k_svd
    small_svd_ = !(ss && ss[0] == '0');
    // knob: 1 runs all groups' SVDs of the grouped executor in lockstep on the
    // grid barrier instead of each group's update on the group's own barrier
    // (same results; slower on the n = 128, chi = 256 chain)
    const char* ls = std::getenv("TESSERA_EXEC_LOCKSTEP");
    ex_group_sync_ = !(ls && ls[0] == '1');
    // knob: 0 keeps the per-group layout on a full barrier after every round
    const char* pp = std::getenv("TESSERA_EXEC_P2P");
    ex_p2p_ = !(pp && pp[0] == '0');
    // knob: 0 adds a grid barrier after every super-step of the per-group layout;
    // by default its updates run in the dataflow order (same results; 9% faster on
    // the n = 128 chain and on the batched sample)
    const char* df = std::getenv("TESSERA_EXEC_DATAFLOW");
    ex_dataflow_ = df && df[0] == '0' ? 0 : 1;
    // knob: 0 gives group g slot g of every super-step in the dataflow order instead of
    // the next update from a shared queue (same results)
    const char* qu = std::getenv("TESSERA_EXEC_QUEUE");
    ex_queue_ = !(qu && qu[0] == '0');
    // knob: 64-cycle sleeps between the polls of a group barrier or round wait
    const char* nap = std::getenv("TESSERA_EXEC_POLL_NAPS");
    ex_naps_ = nap ? std::max(0, std::atoi(nap)) : 1;
    // knob: the Jacobi sweeps of the executor's SVDs end after a sweep whose
    // rotations all had a relative coupling of at most this value (0: after a
    // sweep that rotates nothing)
    const char* st = std::getenv("TESSERA_SVD_STOP");
    svd_stop_ = st ? std::max(0.0, std::atof(st)) : 0.0;
    // knob: 0 runs the executor's tournament rounds with the columns in LDS
    // instead of the row layout
    const char* rw = std::getenv("TESSERA_SVD_ROWS");
    svd_rows_ = !(rw && rw[0] == '0');
    // knob: 0 keeps the executor's SVDs from ending with the matrix-core convergence
    // check (they then end after a sweep that rotates nothing)
    const char* gm = std::getenv("TESSERA_SVD_GRAM");
    svd_gram_ = !(gm && gm[0] == '0');
    // knob: 1 adds the cycle profile of the row-layout rotation steps to the
    // executor profile (its timer reads slow the profiled workgroup)
    const char* sp = std::getenv("TESSERA_EXEC_STEP_PROFILE");
    ex_step_prof_ = sp && sp[0] == '1';
    // knob: 0 switches workgroup 0's phase profile off (its timers cost that
    // workgroup a little, and every round waits for its slowest worker)
    const char* po = std::getenv("TESSERA_EXEC_PROFILE");
    ex_prof_on_ = !(po && po[0] == '0');
    // test knob: polls before a group barrier wait gives up (forces the error path)
    const char* sl = std::getenv("TESSERA_EXEC_SPIN_LIMIT");
    ex_spin_limit_ = sl ? std::max(0ll, std::atoll(sl)) : kSpinLimit;
    // knob, off by default: 1 synchronises each group of the grouped executor
    // XCD-locally when the groups sit one per XCD. Its first run at n = 128,
    // chi = 256 left the GPU faulted (smaller runs were exact); the cause is not
    // established, so the default keeps device-scope barriers.
    const char* gr = std::getenv("TESSERA_STEP_GRAPH");  // knob: 0 launches every update of a step directly
    graph_on_ = !(gr && gr[0] == '0');
    const char* xl = std::getenv("TESSERA_EXEC_XCD_LOCAL");
    ex_xcd_local_ = xl && xl[0] == '1';
    const char* xp = std::getenv("TESSERA_SVD_XCD_PACK");  // 1: keep the SVD's workgroups on one XCD
    xcd_stride_ = (xp && xp[0] == '1') ? 8 : 1;
    build_plan();
    reset_product();
}

Tebd::~Tebd() {
    drop_graph();
    for (auto& e : events_) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
}

// mps.cpp:126-145 routing of one ZZ slice in edge-list order.
void Tebd::build_plan() {
    plan_.clear();
    spans_.clear();
    for (std::size_t e = 0; e < model_.ew.size(); ++e) {
        int i = model_.ei[e], j = model_.ej[e];
        if (i > j) std::swap(i, j);
        if (i == j) continue;
        for (int k = j; k > i + 1; --k) plan_.push_back({k - 1, -1});
        plan_.push_back({i, (int)e});
        for (int k = i + 1; k < j; ++k) plan_.push_back({k, -1});
        spans_.push_back({plan_.size() - (std::size_t)(2 * (j - i) - 1), plan_.size()});
    }
    // a step can be captured into a graph when no update needs the cooperative
    // (multi-workgroup) SVD launch; the choice below mirrors two_site's
    plan_graphable_ = true;
    for (const Update& u : plan_) {
        const int Dlc = cap_[u.site], Drc = cap_[u.site + 2];
        const std::size_t bytes = (std::size_t)(2 * Dlc) * (2 * Drc) * sizeof(double2);
        const int ccap = std::min(2 * Dlc, 2 * Drc), mcap = std::max(2 * Dlc, 2 * Drc);
        const bool small = small_svd_ && !force_dist_ && ccap <= kSmall && mcap <= kSmallTall;
        const bool wg = bytes <= 144 * 1024 && !force_dist_;
        plan_graphable_ = plan_graphable_ && (small || wg);
    }
}

void Tebd::ensure_trace(int count) {
    if (count > trace_cap_) {
        DevBuf<double> nb(count);
        if (trace_cap_) TQ_HIP(hipMemcpyAsync(nb.p, trace_.p, trace_cap_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        TQ_HIP(hipStreamSynchronize(stream_));
        trace_ = std::move(nb);
        trace_cap_ = count;
    }
}

// Largest interior bond dims[1..n-1] (at least 1) into *out: one wavefront.
// In a replayed step (par non-null) the slot is par[2] + par[0].
__global__ void k_max_bond(const int* __restrict__ dims, int n, int* __restrict__ out, const int* par) {
    int m = 1;
    for (int i = 1 + (int)threadIdx.x; i < n; i += 64) m = max(m, dims[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if (threadIdx.x == 0) out[par ? par[2] + par[0] : 0] = m;
}

void Tebd::ensure_max_bond(int count) {
    if (count > mbond_cap_) {
        TQ_HIP(hipStreamSynchronize(stream_));
        mbond_ = DevBuf<int>(count);
        mbond_cap_ = count;
    }
}

void Tebd::record_max_bond(int slot) {
    if (slot < 0 || slot >= mbond_cap_) throw Error(-1, "max-bond slot out of range");
    hipLaunchKernelGGL(k_max_bond, dim3(1), dim3(64), 0, stream_, dims_.p, n_, capturing_ ? mbond_.p : mbond_.p + slot,
                       capturing_ ? sp_.par : nullptr);
    TQ_LAUNCH_CHECK();
}

std::vector<int> Tebd::max_bonds(int count) {
    if (count > mbond_cap_) throw Error(-1, "max-bond count out of range");
    std::vector<int> out(count);
    if (count) TQ_HIP(hipMemcpyAsync(out.data(), mbond_.p, count * sizeof(int), hipMemcpyDeviceToHost, stream_));
    sync();
    return out;
}

void Tebd::reset_product(const double* warm) {
    std::vector<double2> host(off_[n_], make_double2(0, 0));
    const double amp = 1.0 / std::sqrt(2.0);
    for (int i = 0; i < n_; ++i) {
        double a0 = amp, a1 = amp;
        if (warm) {
            // ground state of -(X + b Z): (1, r - b) / norm with r = sqrt(1 + b^2)
            const double b = warm[i], r = std::sqrt(1.0 + b * b);
            const double c1 = b >= 0.0 ? 1.0 / (r + b) : r - b;
            const double nrm = std::sqrt(1.0 + c1 * c1);
            a0 = 1.0 / nrm;
            a1 = c1 / nrm;
        }
        host[off_[i] + 0 * cap_[i + 1]] = make_double2(a0, 0);  // (0,0,0)
        host[off_[i] + 1 * cap_[i + 1]] = make_double2(a1, 0);  // (0,1,0)
    }
    TQ_HIP(hipMemcpyAsync(sites_.p, host.data(), host.size() * sizeof(double2), hipMemcpyHostToDevice, stream_));
    std::vector<int> ones(n_ + 1, 1);
    TQ_HIP(hipMemcpyAsync(dims_.p, ones.data(), ones.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    TQ_HIP(hipMemsetAsync(trunc_.p, 0, kStatSlots * sizeof(double), stream_));
    TQ_HIP(hipStreamSynchronize(stream_));
}

void Tebd::randomize(uint64_t seed, double decay) {
    int maxsite = 1;
    for (int i = 0; i < n_; ++i) maxsite = std::max(maxsite, cap_[i] * 2 * cap_[i + 1]);
    dim3 grid((unsigned)std::min<long long>((maxsite + 255) / 256, 1024), n_);
    hipLaunchKernelGGL(k_randomize, grid, dim3(256), 0, stream_, sites_.p, off_dev_.p, caps_dev_.p, n_,
                       (unsigned long long)seed, decay);
    TQ_LAUNCH_CHECK();
    TQ_HIP(hipMemcpyAsync(dims_.p, cap_.data(), (n_ + 1) * sizeof(int), hipMemcpyHostToDevice, stream_));
    // one left-to-right sweep of identity updates: every site but the last
    // becomes an isometry and bond b carries the weights w_k ~ exp(-k/decay)
    std::vector<double> w(maxcap_);
    double tot = 0.0;
    for (int k = 0; k < maxcap_; ++k) tot += (w[k] = std::exp(-2.0 * k / decay));
    for (int k = 0; k < maxcap_; ++k) w[k] = std::sqrt(w[k] / tot);
    DevBuf<double> wd(maxcap_);
    TQ_HIP(hipMemcpyAsync(wd.p, w.data(), maxcap_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    Gate4 eye;
    for (auto& x : eye.g) x = make_double2(0, 0);
    for (int k = 0; k < 4; ++k) eye.g[k * 5] = make_double2(1, 0);
    const bool timing = timing_;
    timing_ = false;
    for (int p = 0; p + 1 < n_; ++p) two_site(p, eye, wd.p);
    timing_ = timing;
    TQ_HIP(hipMemsetAsync(trunc_.p, 0, kStatSlots * sizeof(double), stream_));
    TQ_HIP(hipStreamSynchronize(stream_));
}

void Tebd::two_site(int p, const Gate4& g, const double* wts, int edge) {
    const int ld = ld_;
    const int Dlc = cap_[p], Drc = cap_[p + 2];
    const long long ttiles = theta_tiles_for(Dlc, Drc);  // one wave per tile at full capacity
    hipLaunchKernelGGL(k_theta, dim3((unsigned)((ttiles + 3) / 4)), dim3(256), 0, stream_, sites_.p, off_dev_.p,
                       caps_dev_.p, dims_.p, p, g, mat_.p, work_.p, ld, sp_, edge);
    TQ_LAUNCH_CHECK();
    // worst-case matrix at this bond decides the kernel (both dims <= 2*cap)
    const std::size_t bytes = (std::size_t)(2 * Dlc) * (2 * Drc) * sizeof(double2);
    if (timing_) {
        if (ev_next_ >= events_.size()) {
            hipEvent_t a, b;
            TQ_HIP(hipEventCreate(&a));
            TQ_HIP(hipEventCreate(&b));
            events_.push_back({a, b});
        }
        TQ_HIP(hipEventRecord(events_[ev_next_].first, stream_));
    }
    const int ccap = std::min(2 * Dlc, 2 * Drc), mcap = std::max(2 * Dlc, 2 * Drc);
    if (small_svd_ && !force_dist_ && ccap <= kSmall && mcap <= kSmallTall) {
        // every column pair of a round on its own 16-lane group, matrix in LDS
        hipLaunchKernelGGL(k_svd_small, dim3(1), dim3(kDistThreads), bytes, stream_, work_.p, ld, dims_.p, p, chi_,
                           sv_.p, perm_.p, trunc_.p, stats_.p);
        TQ_LAUNCH_CHECK();
    } else if (bytes <= 144 * 1024 && !force_dist_) {
        hipLaunchKernelGGL(k_svd, dim3(1), dim3(kSvdThreads), bytes, stream_, work_.p, ld, dims_.p, p, chi_, sv_.p,
                           perm_.p, trunc_.p, 1, stats_.p);
        TQ_LAUNCH_CHECK();
    } else {
        const int nb = (ccap + kBlk - 1) / kBlk;
        const int nbp = std::max(2, nb + (nb & 1));
        const unsigned nwg = (unsigned)(nbp / 2) * (unsigned)xcd_stride_;
        // the LDS columns of a round, or the row layout's reduction area
        const std::size_t shm = std::max({(std::size_t)2 * kBlk * mcap * sizeof(double2), sizeof(RowLds),
                                          lq_lds_bytes(ccap)});
        int* flags = flags_.p;
        int pp = p, cc = chi_;
        unsigned long long* stats = stats_.p;
        int xs = xcd_stride_;
        double2* lqa = svd_lq_ ? aux_.p : nullptr;  // no reflector buffer: no LQ preconditioning
        void* args[] = {(void*)&work_.p, (void*)&ld_,   (void*)&dims_.p,  (void*)&pp,   (void*)&cc,
                        (void*)&sv_.p,   (void*)&perm_.p, (void*)&trunc_.p, (void*)&flags, (void*)&stats,
                        (void*)&xs,      (void*)&lqa,     (void*)&taus_.p};
        TQ_HIP(hipLaunchCooperativeKernel((const void*)k_svd_dist, dim3(nwg), dim3(kDistThreads), args, shm, stream_));
    }
    if (timing_) TQ_HIP(hipEventRecord(events_[ev_next_++].second, stream_));
    const long long work = (long long)(2 * Dlc) * std::min(chi_, 2 * Drc) + (long long)std::min(chi_, 2 * Dlc) * (2 * Drc);
    hipLaunchKernelGGL(k_recon, dim3((unsigned)grid_for(work, 256)), dim3(256), 0, stream_, sites_.p, off_dev_.p,
                       caps_dev_.p, dims_.p, p, mat_.p, work_.p, ld, sv_.p, perm_.p, wts);
    TQ_LAUNCH_CHECK();
}

static Gate4 make_swap_gate() {
    Gate4 g;
    for (auto& x : g.g) x = make_double2(0, 0);
    g.g[0 * 4 + 0] = make_double2(1, 0);
    g.g[1 * 4 + 2] = make_double2(1, 0);
    g.g[2 * 4 + 1] = make_double2(1, 0);
    g.g[3 * 4 + 3] = make_double2(1, 0);
    return g;
}

// exp(-i theta Z(x)Z) = diag(e^{-i th}, e^{i th}, e^{i th}, e^{-i th})  (mps.cpp:116-124)
static Gate4 make_zz_gate(double theta) {
    Gate4 g;
    for (auto& x : g.g) x = make_double2(0, 0);
    const double c = std::cos(theta), s = std::sin(theta);
    g.g[0] = make_double2(c, -s);
    g.g[5] = make_double2(c, s);
    g.g[10] = make_double2(c, s);
    g.g[15] = make_double2(c, -s);
    return g;
}

void Tebd::step(double A, double B, int rec) {
    const double* bias = bias_.p;
    hipLaunchKernelGGL(k_single_layer, dim3(n_), dim3(256), 0, stream_, sites_.p, off_dev_.p, caps_dev_.p, dims_.p,
                       h_.p, bias, A, B, dt_, sp_);
    TQ_LAUNCH_CHECK();
    static const Gate4 sw = make_swap_gate();
    const double sub = dt_ / ns_;
    auto run_edge = [&](const std::pair<std::size_t, std::size_t>& span) {
        for (std::size_t q = span.first; q < span.second; ++q) {
            const Update& u = plan_[q];
            two_site(u.site, u.edge < 0 ? sw : make_zz_gate(-B * model_.ew[u.edge] * sub), nullptr, u.edge);
        }
    };
    for (int s = 0; s < ns_; ++s) {
        // odd slices visit the edge list backwards, each edge's own swap-in /
        // gate / swap-out sequence unchanged (mps.cpp:333-340)
        if (s % 2 == 0)
            for (const auto& span : spans_) run_edge(span);
        else
            for (std::size_t k = spans_.size(); k-- > 0;) run_edge(spans_[k]);
    }
    hipLaunchKernelGGL(k_single_layer, dim3(n_), dim3(256), 0, stream_, sites_.p, off_dev_.p, caps_dev_.p, dims_.p,
                       h_.p, bias, A, B, dt_, sp_);
    TQ_LAUNCH_CHECK();
    if (rec >= 0) entropy_sweep(rec);
}

void Tebd::run(const double* a, const double* b, int count, int rec0, int bond0) {
    if (count <= 0) return;
    if (rec0 >= 0) ensure_trace(rec0 + count);
    if (bond0 >= 0) ensure_max_bond(bond0 + count);
    if (exec_req_ != 0 && n_ >= 2 && (bond0 < 0 || rec0 >= 0)) {
        // snapshots of a recorded chunk stay resident until its entropies are
        // taken; bound the chunk to ~8 GiB of snapshots
        const std::size_t per = off_[n_] * sizeof(double2);
        const int chunk = rec0 >= 0 ? (int)std::max<std::size_t>(1, std::min<std::size_t>(64, (8ull << 30) / per)) : 64;
        for (int k0 = 0; k0 < count; k0 += chunk) {
            const int c = std::min(chunk, count - k0);
            run_exec(a + k0, b + k0, c, rec0 >= 0 ? rec0 + k0 : -1, bond0 >= 0 ? bond0 + k0 : -1);
        }
        return;
    }
    if (graph_on_ && !timing_ && plan_graphable_ && n_ >= 2 && count >= 2) {
        run_graph(a, b, count, rec0, bond0);
        return;
    }
    for (int k = 0; k < count; ++k) {
        step(a[k], b[k], rec0 >= 0 ? rec0 + k : -1);
        if (bond0 >= 0) record_max_bond(bond0 + k);
    }
}

// Per-update path as a hipGraph: one Trotter step's launch sequence (single-site
// layers, every two-site update of the routed ZZ block, the entropy sweep when
// recording, and a node advancing the step index) is captured once and
// replayed `count` times. Its kernels read the step's schedule values, ZZ gate
// phasors and entropy slot from device memory (StepParams), computed on the host
// with the same calls as the launch-per-update path, so both give the same bits.
void Tebd::run_graph(const double* a, const double* b, int count, int rec0, int bond0) {
    sync();  // the previous run's host staging below is free again
    const int m = (int)model_.ew.size();
    const double sub = dt_ / ns_;
    g_host_ab_.resize(2 * (std::size_t)count);
    g_host_zz_.resize((std::size_t)count * std::max(m, 1));
    for (int k = 0; k < count; ++k) {
        g_host_ab_[2 * k] = a[k];
        g_host_ab_[2 * k + 1] = b[k];
        for (int e = 0; e < m; ++e) {  // make_zz_gate(-B J dt / ns)
            const double th = -b[k] * model_.ew[e] * sub;
            g_host_zz_[(std::size_t)k * m + e] = make_double2(std::cos(th), std::sin(th));
        }
    }
    if (g_ab_.n < g_host_ab_.size()) g_ab_.alloc(g_host_ab_.size());
    if (g_zz_.n < g_host_zz_.size()) g_zz_.alloc(g_host_zz_.size());
    if (!g_par_.p) g_par_.alloc(3);
    g_host_par_[0] = 0;
    g_host_par_[1] = rec0 >= 0 ? rec0 : 0;
    g_host_par_[2] = bond0 >= 0 ? bond0 : 0;
    TQ_HIP(hipMemcpyAsync(g_ab_.p, g_host_ab_.data(), g_host_ab_.size() * sizeof(double), hipMemcpyHostToDevice,
                          stream_));
    TQ_HIP(hipMemcpyAsync(g_zz_.p, g_host_zz_.data(), g_host_zz_.size() * sizeof(double2), hipMemcpyHostToDevice,
                          stream_));
    TQ_HIP(hipMemcpyAsync(g_par_.p, g_host_par_, 3 * sizeof(int), hipMemcpyHostToDevice, stream_));
    const bool record = rec0 >= 0, bonds = bond0 >= 0;
    const int mode = (int)record | ((int)bonds << 1);
    const void* key[5] = {g_ab_.p, g_zz_.p, trace_.p, bias_.p, mbond_.p};
    bool same = graph_exec_ && graph_rec_ == mode;
    for (int q = 0; q < 5; ++q) same = same && graph_key_[q] == key[q];
    if (!same) {
        drop_graph();
        sp_ = StepParams{g_ab_.p, g_zz_.p, g_par_.p, m};
        capturing_ = true;
        TQ_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            step(0.0, 0.0, record ? 0 : -1);  // A, B and the ZZ gates come from StepParams
            if (bonds) record_max_bond(0);
            hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, stream_, g_par_.p);
            TQ_LAUNCH_CHECK();
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            capturing_ = false;
            sp_ = StepParams{};
            throw;
        }
        capturing_ = false;
        sp_ = StepParams{};
        TQ_HIP(hipStreamEndCapture(stream_, &g));
        graph_ = g;
        TQ_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_rec_ = mode;
        for (int q = 0; q < 5; ++q) graph_key_[q] = key[q];
    }
    for (int k = 0; k < count; ++k) TQ_HIP(hipGraphLaunch(graph_exec_, stream_));
}

void Tebd::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_rec_ = -1;
}

void Tebd::set_svd_lq(bool on) {
    sync();
    svd_lq_ = on;
}

void Tebd::set_executor(int groups, int workers) {
    if (groups < -1) throw Error(-1, "executor groups must be >= -1");
    if (workers < 0) throw Error(-1, "executor workers must be >= 0");
    sync();
    exec_req_ = groups;
    exec_workers_ = workers;
}

// Groups of `per_group` workgroups (enough for the widest block-Jacobi round of
// a theta with rows and columns up to 2*maxcap), as many groups as are
// co-resident (at most `req` if req > 0, else 64), a multiple of 8 when 8 or
// more fit.
static const void* exec_kernel(bool solo, bool group_sync, bool xl = false) {
    if (solo) return (const void*)k_tebd_lockstep<kExecSolo>;
    if (group_sync)
        return xl ? (const void*)k_tebd_lockstep<kExecGroupSync, true> : (const void*)k_tebd_lockstep<kExecGroupSync>;
    return (const void*)k_tebd_lockstep<kExecLockstep>;
}

// Workgroup i of a launch runs on XCD i mod 8 (the dispatcher deals workgroups
// to the XCDs round-robin): checked once per device with a probe launch that
// reads each workgroup's XCC id. The executor's XCD-local layout relies on it
// (and re-checks it in every launch).
__global__ void k_xcc_map(int* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = (int)xcc_id();
}
static bool xcd_round_robin() {
    static std::once_flag once[64];
    static bool ok[64] = {};
    int dev = 0;
    TQ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return false;
    std::call_once(once[dev], [&] {
        constexpr int nb = 1024;
        DevBuf<int> d(nb);
        hipLaunchKernelGGL(k_xcc_map, dim3(nb), dim3(64), 0, nullptr, d.p);
        if (hipGetLastError() != hipSuccess) return;
        std::vector<int> h(nb);
        if (hipMemcpy(h.data(), d.p, nb * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return;
        bool good = true;
        for (int b = 0; b < nb; ++b) good = good && h[b] == (b & 7);
        ok[dev] = good;
    });
    return ok[dev];
}

static void exec_shape(int maxcap, int req, bool solo, bool group_sync, int workers, int& groups, int& per_group,
                       std::size_t& shm, bool xl = false) {
    const int cmax = 2 * maxcap;
    const int nb = (cmax + kBlk - 1) / kBlk;
    const int nbp = std::max(2, nb + (nb & 1));
    // grouped: one worker per block pair of a tournament round by default; fewer
    // workers take several pairs each (fewer workgroups in every barrier)
    per_group = solo ? 1 : (workers > 0 ? std::min(workers, nbp / 2) : nbp / 2);
    const std::size_t svd_bytes = (std::size_t)2 * kBlk * (2 * maxcap) * sizeof(double2);
    // solo: room for a kSmall x kSmall theta of svd_small
    const std::size_t small_bytes = solo ? (std::size_t)kSmall * kSmall * sizeof(double2) : 0;
    shm = std::max({svd_bytes, small_bytes, sizeof(RowLds), lq_lds_bytes(cmax)});
    int per_cu = 0, cus = 0, dev = 0;
    TQ_HIP(hipGetDevice(&dev));
    TQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    TQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, exec_kernel(solo, group_sync, xl),
                                                        exec_threads(solo ? kExecSolo : kExecLockstep), shm));
    const long long resident = std::min<long long>((long long)per_cu * cus, kStatSlots);
    long long g = resident / per_group;
    const long long cap = req > 0 ? req : (solo ? kStatSlots : 64);
    g = std::min(g, cap);
    if (g >= 8) g -= g % 8;
    if (g < 1) throw Error(-3, "lockstep executor: not enough co-resident workgroups");
    groups = (int)g;
}

void Tebd::run_exec(const double* a, const double* b, int count, int rec0, int bond0) {
    std::vector<Tebd*> one{this};
    run_lockstep(one, {a}, {b}, {count}, rec0, exec_req_, 1, bond0);
}

// The lockstep executor over one or several MPS (batched runs): every
// instance's updates are placed, in its own gate order, into super-steps of at
// most G mutually independent updates shared by all instances.
void Tebd::run_lockstep(const std::vector<Tebd*>& ts, const std::vector<const double*>& as,
                        const std::vector<const double*>& bs, const std::vector<int>& counts, int rec0, int req,
                        int mode, int bond0) {
    if (ts.empty()) return;
    if (rec0 >= 0 && ts.size() != 1) throw Error(-1, "entropy recording needs a single instance");
    if (mode < 0 || mode > 2) throw Error(-1, "executor mode must be 0 (auto), 1 (grouped) or 2 (solo)");
    Tebd& t0 = *ts[0];
    for (Tebd* t : ts) TQ_HIP(hipStreamSynchronize(t->stream_));  // previous runs' buffers are free again
    int maxcap = 1;
    for (Tebd* t : ts) maxcap = std::max(maxcap, t->maxcap_);
    const int ld = 2 * maxcap;
    // Grouped: each update's SVD spread over a group of workgroups in lockstep
    // (few independent updates, large theta). Solo: one workgroup per update
    // and one grid barrier per super-step; thetas up to 64 x 64 stay in LDS
    // (svd_small). Automatic: solo for batches of four or more instances whose
    // bonds stay at or below 32, where every theta is small.
    const bool solo = mode == 2 || (mode == 0 && ts.size() >= 4 && maxcap <= 32);
    int G = 0, W = 0;
    std::size_t shm = 0;
    const bool group_sync = !solo && t0.ex_group_sync_;
    // dataflow order (per-group layout, no super-step barriers): every snapshot is
    // taken by the latest update on its site, or before any update when there is none
    const bool dataflow = group_sync && t0.ex_dataflow_ != 0;
    // XCD-local groups: grouped layout, one group per XCD (G a multiple of 8),
    // not in the dataflow order (whose done flags cross XCDs)
    bool xl = group_sync && t0.ex_xcd_local_ && !dataflow && xcd_round_robin();
    int Gxl = 0, Wxl = 0;
    std::size_t shm_xl = 0;
    if (xl) {
        exec_shape(maxcap, req, solo, group_sync, ts[0]->exec_workers_, Gxl, Wxl, shm_xl, true);
        xl = Gxl >= 8 && Gxl % 8 == 0;
    }
    exec_shape(maxcap, req, solo, group_sync, ts[0]->exec_workers_, G, W, shm);
    using cd = std::complex<double>;
    struct M2 {
        cd u[2][2];
    };
    const M2 eye{{{1.0, 0.0}, {0.0, 1.0}}};
    auto mul2 = [](const M2& x, const M2& y) {
        M2 r;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) r.u[i][j] = x.u[i][0] * y.u[0][j] + x.u[i][1] * y.u[1][j];
        return r;
    };
    auto to_d2 = [](cd z) { return make_double2(z.real(), z.imag()); };
    auto from_d2 = [](double2 z) { return cd(z.x, z.y); };
    const Gate4 sw = make_swap_gate();
    std::vector<std::vector<ExecUpdate>> ups;  // per super-step, at most G
    // per update (parallel to ups): (super-step << 16 | slot) of its predecessors on
    // its two sites, -1: none; flat update indices once the offsets are known
    std::vector<std::vector<std::array<long long, 2>>> pred_pos;
    std::vector<int> rounds;
    struct SnapAt {
        int at;
        long long pos;  // (super-step << 16 | slot) of the site's latest update, -1: none
        ExecSnap s;
    };
    std::vector<SnapAt> snaps;
    int nsteps = 0;
    std::vector<std::vector<M2>> pendings(ts.size());
    std::vector<int> last_ss(ts.size(), -1);  // last super-step holding an update of each instance
    // open[t]: a super-step >= t with a free slot (path-compressed), so that
    // placing the updates of many instances stays linear in their number
    std::vector<int> open;
    auto first_open = [&](int t) {
        int r = t;
        while (r < (int)open.size() && open[r] != r) r = open[r];
        while (t < (int)open.size() && open[t] != t) {
            const int nx = open[t];
            open[t] = r;
            t = nx;
        }
        return r;
    };
    for (std::size_t in = 0; in < ts.size(); ++in) {
        Tebd& T = *ts[in];
        const int n = T.n_;
        const double* a = as[in];
        const double* b = bs[in];
        // exp(-i dt/2 g), g = [[-hz, -A], [-A, hz]] (the closed form of k_single_layer)
        auto half_layer = [&](double A, double B, int i) {
            const double hz = B * T.model_.h[i] + (T.model_.bias.empty() ? 0.0 : A * T.model_.bias[i]);
            const double r = std::sqrt(A * A + hz * hz);
            M2 u = eye;
            if (r > 0.0) {
                const double sn = std::sin(0.5 * T.dt_ * r), cs = std::cos(0.5 * T.dt_ * r);
                u.u[0][0] = cd(cs, sn * hz / r);
                u.u[0][1] = u.u[1][0] = cd(0.0, sn * A / r);
                u.u[1][1] = cd(cs, -sn * hz / r);
            }
            return u;
        };
        auto nbp_of = [&](int p) {  // tournament length of the widest matrix update p can see
            const int c = std::min(2 * T.cap_[p], 2 * T.cap_[p + 2]);
            const int nb = (c + kBlk - 1) / kBlk;
            return std::max(2, nb + (nb & 1));
        };
        std::vector<M2>& pending = pendings[in];
        pending.assign(n, eye);
        last_ss[in] = -1;
        std::vector<int> free_at(n, 0);   // first super-step at which the site may be updated
        std::vector<int> last_at(n, -1);  // super-step of the site's latest update
        // (super-step << 16 | slot) of the site's latest update, -1: none; converted
        // to flat update indices once every super-step's offset is known
        std::vector<long long> last_pos(n, -1);
        auto place = [&](int p, const Gate4& g) {
            const int t = first_open(std::max(free_at[p], free_at[p + 1]));
            if (t >= (int)ups.size()) {
                ups.resize(t + 1);
                pred_pos.resize(t + 1);
                rounds.resize(t + 1, 0);
                for (int q = (int)open.size(); q <= t; ++q) open.push_back(q);
            }
            ExecUpdate u;
            u.site = p;
            u.inst = (int)in;
            u.g = g;
            const long long pa = last_pos[p], pb = last_pos[p + 1];
            u.pred[0] = u.pred[1] = -1;
            pred_pos[t].push_back({pa, pb == pa ? -1 : pb});
            last_pos[p] = last_pos[p + 1] = ((long long)t << 16) | (long long)ups[t].size();
            ups[t].push_back(u);
            if ((int)ups[t].size() >= G) open[t] = t + 1;  // full: later searches skip it
            rounds[t] = std::max(rounds[t], nbp_of(p) - 1);
            last_ss[in] = std::max(last_ss[in], t);
            free_at[p] = free_at[p + 1] = t + 1;
            nsteps = std::max(nsteps, t + 1);
            return t;
        };
        for (int k = 0; k < counts[in]; ++k) {
            const double A = a[k], B = b[k];
            std::vector<M2> u1(n);
            for (int i = 0; i < n; ++i) {
                u1[i] = half_layer(A, B, i);
                pending[i] = mul2(u1[i], pending[i]);
            }
            const double sub = T.dt_ / T.ns_;
            auto add_edge = [&](const std::pair<std::size_t, std::size_t>& span) {
                for (std::size_t q = span.first; q < span.second; ++q) {
                    const Update& u = T.plan_[q];
                    const Gate4 g = u.edge < 0 ? sw : make_zz_gate(-B * T.model_.ew[u.edge] * sub);
                    const M2& P = pending[u.site];
                    const M2& Q = pending[u.site + 1];
                    Gate4 f;
                    for (int r = 0; r < 4; ++r)
                        for (int c = 0; c < 4; ++c) {
                            cd acc = 0.0;
                            for (int s = 0; s < 4; ++s)
                                acc += from_d2(g.g[r * 4 + s]) * P.u[s >> 1][c >> 1] * Q.u[s & 1][c & 1];
                            f.g[r * 4 + c] = to_d2(acc);
                        }
                    pending[u.site] = eye;
                    pending[u.site + 1] = eye;
                    const int t = place(u.site, f);
                    last_at[u.site] = last_at[u.site + 1] = t;
                }
            };
            for (int s = 0; s < T.ns_; ++s) {
                if (s % 2 == 0)
                    for (const auto& span : T.spans_) add_edge(span);
                else
                    for (std::size_t e = T.spans_.size(); e-- > 0;) add_edge(T.spans_[e]);
            }
            for (int i = 0; i < n; ++i) pending[i] = mul2(u1[i], pending[i]);
            if (rec0 >= 0)
                for (int i = 0; i < n; ++i) {
                    // after the site's last update of this step (sites change only in
                    // the reconstruction phase), before any later update changes it
                    SnapAt x;
                    x.at = last_at[i] + 1;
                    x.pos = last_pos[i];
                    x.s.site = i;
                    x.s.slot = k;
                    x.s.g[0] = to_d2(pending[i].u[0][0]);
                    x.s.g[1] = to_d2(pending[i].u[0][1]);
                    x.s.g[2] = to_d2(pending[i].u[1][0]);
                    x.s.g[3] = to_d2(pending[i].u[1][1]);
                    snaps.push_back(x);
                }
        }
    }
    // snapshots due after the last super-step get a trailing super-step of their own
    int total_steps = nsteps;
    for (const auto& x : snaps) total_steps = std::max(total_steps, x.at + 1);
    if (total_steps > (int)rounds.size()) {
        ups.resize(total_steps);
        rounds.resize(total_steps, 0);
    }
    std::vector<int> up_begin(total_steps + 1, 0);
    int widest = 1;
    for (int t = 0; t < total_steps; ++t) {
        up_begin[t + 1] = up_begin[t] + (int)ups[t].size();
        widest = std::max(widest, (int)ups[t].size());
    }
    G = std::min(G, widest);  // launch only as many groups as a super-step uses
    if (xl) {  // ... rounded up to whole XCDs
        G = std::min(Gxl, (widest + 7) / 8 * 8);
        W = Wxl;
        shm = shm_xl;
    }
    const int nup = up_begin[total_steps];
    std::vector<int> snap_begin(total_steps + 1, 0);
    std::vector<int> usnap_begin;  // dataflow: [nup + 1]
    int npro = 0;
    std::vector<ExecSnap> snap_list(snaps.size());
    if (dataflow) {
        auto owner = [&](const SnapAt& x) {  // flat index of the update taking it, -1: before any
            return x.pos < 0 ? -1 : up_begin[(int)(x.pos >> 16)] + (int)(x.pos & 0xffff);
        };
        std::stable_sort(snaps.begin(), snaps.end(),
                         [&](const SnapAt& x, const SnapAt& y) { return owner(x) < owner(y); });
        usnap_begin.assign((std::size_t)nup + 1, 0);
        for (std::size_t q = 0; q < snaps.size(); ++q) {
            snap_list[q] = snaps[q].s;
            const int f = owner(snaps[q]);
            if (f < 0)
                ++npro;
            else
                ++usnap_begin[f + 1];
        }
        usnap_begin[0] = npro;
        for (int f = 0; f < nup; ++f) usnap_begin[f + 1] += usnap_begin[f];
    } else {
        std::stable_sort(snaps.begin(), snaps.end(), [](const SnapAt& x, const SnapAt& y) { return x.at < y.at; });
        for (std::size_t q = 0; q < snaps.size(); ++q) {
            snap_list[q] = snaps[q].s;
            ++snap_begin[snaps[q].at + 1];
        }
        for (int t = 0; t < total_steps; ++t) snap_begin[t + 1] += snap_begin[t];
    }
    // a super-step whose slots are all idle still runs one (empty) round
    for (int t = 0; t < total_steps; ++t) rounds[t] = std::max(rounds[t], 1);

    // device buffers of instance 0 (grown on demand, kept across runs)
    const std::size_t mat_sz = (std::size_t)ld * ld;
    if (t0.ex_mat_.n < G * mat_sz) {
        t0.ex_mat_.alloc(G * mat_sz);
        t0.ex_work_.alloc(G * mat_sz);
        t0.ex_aux_.alloc(G * mat_sz);
    }
    if (t0.ex_sv_.n < (std::size_t)G * ld) {
        t0.ex_sv_.alloc((std::size_t)G * ld);
        t0.ex_perm_.alloc((std::size_t)G * ld);
        t0.ex_taus_.alloc((std::size_t)G * ld);
    }
    if (t0.ex_flags_.n < 8 * (std::size_t)G + 3) t0.ex_flags_.alloc(8 * (std::size_t)G + 3);
    if (t0.ex_bar_.n < 2 * (std::size_t)G + 1) t0.ex_bar_.alloc(2 * (std::size_t)G + 1);
    if (!t0.ex_prof_.p) {
        t0.ex_prof_.alloc(kProfSlots);
        TQ_HIP(hipMemsetAsync(t0.ex_prof_.p, 0, kProfSlots * sizeof(unsigned long long), t0.stream_));
    }
    const int count0 = counts[0];
    if (rec0 >= 0 && t0.ex_snap_.n < (std::size_t)count0 * t0.off_[t0.n_]) {
        t0.ex_snap_.alloc((std::size_t)count0 * t0.off_[t0.n_]);
        t0.ex_sdims_.alloc((std::size_t)count0 * (t0.n_ + 1));
    }
    std::vector<ExecInst> insts(ts.size());
    for (std::size_t in = 0; in < ts.size(); ++in) {
        Tebd& T = *ts[in];
        insts[in] = ExecInst{T.sites_.p, T.off_dev_.p, T.caps_dev_.p, T.dims_.p, T.chi_, T.trunc_.p, T.stats_.p};
    }
    auto align = [](std::size_t x) { return (x + 255) & ~(std::size_t)255; };
    const std::size_t in_bytes = align(insts.size() * sizeof(ExecInst));
    const std::size_t up_bytes = align(std::max<std::size_t>(1, up_begin[total_steps]) * sizeof(ExecUpdate));
    const std::size_t sn_bytes = align(std::max<std::size_t>(1, snap_list.size()) * sizeof(ExecSnap));
    for (int t = 0; t < nsteps; ++t)  // predecessors as flat update indices
        for (std::size_t k = 0; k < ups[t].size(); ++k)
            for (int q = 0; q < 2; ++q) {
                const long long e = pred_pos[t][k][q];
                ups[t][k].pred[q] = e < 0 ? -1 : up_begin[(int)(e >> 16)] + (int)(e & 0xffff);
            }
    // the shared queue's order: by dependency depth (one more than the deepest
    // predecessor's, with no cap on the updates per level), then flat order. The
    // super-steps hold at most G updates, so once the steps in flight offer more
    // than G independent updates they push critical updates behind later steps' ones;
    // the uncapped depth puts them first (a CPU model of 12 steps at n = 128 with 8
    // groups: 35.6 s against 44.6 s in the flat order; the same at 6 steps, where the
    // critical path sets the time). Predecessors are shallower, so the order is
    // topological and every claimed update's predecessors were claimed before it.
    std::vector<int> qorder;
    if (dataflow && t0.ex_queue_) {
        std::vector<int> depth((std::size_t)nup, 0);
        for (int t = 0; t < nsteps; ++t)
            for (std::size_t k = 0; k < ups[t].size(); ++k) {
                const int f = up_begin[t] + (int)k;
                for (int q = 0; q < 2; ++q)
                    if (ups[t][k].pred[q] >= 0) depth[f] = std::max(depth[f], depth[ups[t][k].pred[q]] + 1);
            }
        qorder.resize((std::size_t)nup);
        for (int f = 0; f < nup; ++f) qorder[f] = f;
        std::stable_sort(qorder.begin(), qorder.end(), [&](int x, int y) { return depth[x] < depth[y]; });
    }
    const std::size_t int_count =
        (std::size_t)total_steps + 2 * (total_steps + 1) + usnap_begin.size() + qorder.size();
    t0.ex_host_.assign(in_bytes + up_bytes + sn_bytes + int_count * sizeof(int), 0);
    unsigned char* hb = t0.ex_host_.data();
    std::memcpy(hb, insts.data(), insts.size() * sizeof(ExecInst));
    for (int t = 0; t < total_steps; ++t)
        if (!ups[t].empty())
            std::memcpy(hb + in_bytes + (std::size_t)up_begin[t] * sizeof(ExecUpdate), ups[t].data(),
                        ups[t].size() * sizeof(ExecUpdate));
    if (!snap_list.empty())
        std::memcpy(hb + in_bytes + up_bytes, snap_list.data(), snap_list.size() * sizeof(ExecSnap));
    int* hi = reinterpret_cast<int*>(hb + in_bytes + up_bytes + sn_bytes);
    std::memcpy(hi, rounds.data(), total_steps * sizeof(int));
    std::memcpy(hi + total_steps, snap_begin.data(), (total_steps + 1) * sizeof(int));
    std::memcpy(hi + 2 * total_steps + 1, up_begin.data(), (total_steps + 1) * sizeof(int));
    if (!usnap_begin.empty())
        std::memcpy(hi + 3 * total_steps + 2, usnap_begin.data(), usnap_begin.size() * sizeof(int));
    if (!qorder.empty())
        std::memcpy(hi + 3 * total_steps + 2 + usnap_begin.size(), qorder.data(), qorder.size() * sizeof(int));
    if (t0.ex_tasks_.n < t0.ex_host_.size()) t0.ex_tasks_.alloc(t0.ex_host_.size());
    TQ_HIP(hipMemcpy(t0.ex_tasks_.p, hb, t0.ex_host_.size(), hipMemcpyHostToDevice));

    ExecArgs x{};
    x.insts = reinterpret_cast<const ExecInst*>(t0.ex_tasks_.p);
    x.ups = reinterpret_cast<const ExecUpdate*>(t0.ex_tasks_.p + in_bytes);
    x.snaps = reinterpret_cast<const ExecSnap*>(t0.ex_tasks_.p + in_bytes + up_bytes);
    x.rounds = reinterpret_cast<const int*>(t0.ex_tasks_.p + in_bytes + up_bytes + sn_bytes);
    x.snap_begin = x.rounds + total_steps;
    x.up_begin = x.rounds + 2 * total_steps + 1;
    x.usnap_begin = usnap_begin.empty() ? nullptr : x.rounds + 3 * total_steps + 2;
    x.qorder = qorder.empty() ? nullptr : x.rounds + 3 * total_steps + 2 + usnap_begin.size();
    x.npro = npro;
    x.small = solo ? 1 : 0;
    x.nsteps = total_steps;
    x.G = G;
    x.n = t0.n_;
    x.mat = t0.ex_mat_.p;
    x.work = t0.ex_work_.p;
    x.ld = ld;
    x.sv = t0.ex_sv_.p;
    x.perm = t0.ex_perm_.p;
    x.svd_flags = t0.ex_flags_.p;
    x.gflags = t0.ex_flags_.p + 8 * G;
    x.snap = t0.ex_snap_.p;
    x.snap_stride = t0.off_[t0.n_];
    x.snap_dims = t0.ex_sdims_.p;
    x.aux = t0.svd_lq_ ? t0.ex_aux_.p : nullptr;  // null: Jacobi on theta itself, no LQ preconditioning
    x.taus = t0.ex_taus_.p;
    x.err = t0.ex_bar_.p;  // the error word, then [G][2] barrier words
    x.spin_limit = t0.ex_spin_limit_;
    x.naps = t0.ex_naps_;
    x.stop_ratio = t0.svd_stop_;
    x.rows = t0.svd_rows_ ? 1 : 0;
    x.gram = t0.svd_gram_ ? 1 : 0;
    x.step_prof = t0.ex_step_prof_ ? 1 : 0;
    x.prof_on = t0.ex_prof_on_ ? 1 : 0;
    x.dataflow = dataflow ? 1 : 0;
    x.done = nullptr;
    x.queue = nullptr;
    x.nup = nup;
    if (x.dataflow && t0.ex_queue_ && nup > 0 && !x.qorder) throw Error(-1, "executor: the queue's order is missing");
    if (x.dataflow) {
        // done flags, then the queue's counter and the groups' claims
        const std::size_t nu = std::max<std::size_t>(1, up_begin[total_steps]);
        const std::size_t nq = t0.ex_queue_ ? 1 + (std::size_t)G : 0;
        if (t0.ex_done_.n < nu + nq) t0.ex_done_.alloc(nu + nq);
        TQ_HIP(hipMemsetAsync(t0.ex_done_.p, 0, (nu + nq) * sizeof(int), t0.stream_));
        x.done = t0.ex_done_.p;
        if (t0.ex_queue_) x.queue = t0.ex_done_.p + nu;
    }
    if (group_sync && t0.ex_p2p_) {
        const std::size_t ne = (std::size_t)G * W * 64;
        if (t0.ex_ep_.n < ne) t0.ex_ep_.alloc(ne);
        TQ_HIP(hipMemsetAsync(t0.ex_ep_.p, 0, ne * sizeof(int), t0.stream_));
        x.epochs = t0.ex_ep_.p;
    } else {
        x.epochs = nullptr;
    }
    x.group_bar = t0.ex_bar_.p + 1;
    x.prof = t0.ex_prof_.p;
    x.ss_clock = nullptr;
    x.done_clock = nullptr;
    // completion clocks: per super-step end, or in the dataflow order per update
    const std::size_t nclk = (std::size_t)(x.dataflow ? nup : total_steps) + 1;
    if (t0.ex_clock_.n < nclk) t0.ex_clock_.alloc(nclk);
    (x.dataflow ? x.done_clock : x.ss_clock) = t0.ex_clock_.p;
    void* args[] = {(void*)&x};
    hipStream_t st = t0.stream_;
    // barrier counters and the error word start at zero (ordinary stream work:
    // this launch sequence is never captured into a graph)
    TQ_HIP(hipMemsetAsync(t0.ex_bar_.p, 0, (2 * (std::size_t)G + 1) * sizeof(int), st));
    t0.ex_armed_ = true;
    if (t0.timing_) {
        if (t0.ev_next_ >= t0.events_.size()) {
            hipEvent_t e0, e1;
            TQ_HIP(hipEventCreate(&e0));
            TQ_HIP(hipEventCreate(&e1));
            t0.events_.push_back({e0, e1});
        }
        TQ_HIP(hipEventRecord(t0.events_[t0.ev_next_].first, st));
    }
    if (total_steps > 0)
        TQ_HIP(hipLaunchCooperativeKernel(exec_kernel(solo, group_sync, xl),
                                          dim3((unsigned)(G * W)), dim3(exec_threads(solo ? kExecSolo : kExecLockstep)),
                                          args, shm, st));
    if (t0.timing_) TQ_HIP(hipEventRecord(t0.events_[t0.ev_next_++].second, st));
    t0.exec_supersteps_ += total_steps;
    t0.exec_updates_ += up_begin[total_steps];
    // fold the single-site gates still pending into the sites
    std::vector<std::vector<double2>> gates(ts.size());
    for (std::size_t in = 0; in < ts.size(); ++in) {
        Tebd& T = *ts[in];
        std::vector<double2>& g = gates[in];
        g.resize(4 * (std::size_t)T.n_);
        for (int i = 0; i < T.n_; ++i) {
            const M2& P = pendings[in][i];
            g[4 * i + 0] = to_d2(P.u[0][0]);
            g[4 * i + 1] = to_d2(P.u[0][1]);
            g[4 * i + 2] = to_d2(P.u[1][0]);
            g[4 * i + 3] = to_d2(P.u[1][1]);
        }
        if (!T.ex_gates_.p) T.ex_gates_.alloc(4 * (std::size_t)T.n_);
        TQ_HIP(hipMemcpyAsync(T.ex_gates_.p, g.data(), g.size() * sizeof(double2), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_apply_site_gates, dim3(T.n_), dim3(256), 0, st, T.sites_.p, T.off_dev_.p, T.caps_dev_.p,
                           T.dims_.p, T.ex_gates_.p);
        TQ_LAUNCH_CHECK();
    }
    // entropies of the recorded steps from their snapshots, and the largest bond
    // after each from the snapshots' bond dimensions (every bond's last change in a
    // step precedes both of its sites' snapshots)
    if (rec0 >= 0)
        for (int k = 0; k < count0; ++k) {
            const int* sd = t0.ex_sdims_.p + (std::size_t)k * (t0.n_ + 1);
            t0.entropy_of(t0.ex_snap_.p + (std::size_t)k * t0.off_[t0.n_], sd, rec0 + k);
            if (bond0 >= 0) {
                if (bond0 + k >= t0.mbond_cap_) throw Error(-1, "max-bond slot out of range");
                hipLaunchKernelGGL(k_max_bond, dim3(1), dim3(64), 0, st, sd, t0.n_, t0.mbond_.p + bond0 + k,
                                   nullptr);
                TQ_LAUNCH_CHECK();
            }
        }
    TQ_HIP(hipStreamSynchronize(st));  // host staging buffers go out of scope
    // per instance: device time from the launch start to the end of the super-step
    // holding its last update, or in the dataflow order to its latest done flag
    for (Tebd* t : ts) t->last_run_ms_ = 0.0;
    if (total_steps > 0) {
        std::vector<unsigned long long> clk(nclk);
        TQ_HIP(hipMemcpy(clk.data(), t0.ex_clock_.p, nclk * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        TQ_HIP(hipGetDevice(&dev));
        TQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (x.dataflow) {  // an instance completes at the latest done flag of its updates
            std::vector<unsigned long long> last(ts.size(), clk[nup]);
            for (int t = 0; t < nsteps; ++t)
                for (std::size_t k = 0; k < ups[t].size(); ++k) {
                    unsigned long long& l = last[ups[t][k].inst];
                    l = std::max(l, clk[up_begin[t] + k]);
                }
            for (std::size_t in = 0; in < ts.size(); ++in)
                if (last_ss[in] >= 0 && khz > 0) ts[in]->last_run_ms_ = (double)(last[in] - clk[nup]) / khz;
        } else {
            for (std::size_t in = 0; in < ts.size(); ++in)
                if (last_ss[in] >= 0 && khz > 0)
                    ts[in]->last_run_ms_ = (double)(clk[last_ss[in]] - clk[total_steps]) / khz;
        }
    }
}

void Tebd::env_right(const double2* rin, int j, double2* rout, const double2* sites, const int* dims) {
    if (!sites) sites = sites_.p;
    if (!dims) dims = dims_.p;
    const int ldr = chi_max();
    const long long w1 = (long long)cap_[j] * 2 * cap_[j + 1];
    hipLaunchKernelGGL(k_envr1, dim3((unsigned)grid_for(w1, 256)), dim3(256), 0, stream_, sites, off_dev_.p,
                       caps_dev_.p, dims, j, rin, ldr, envT_.p);
    TQ_LAUNCH_CHECK();
    const long long w2 = (long long)cap_[j] * cap_[j];
    hipLaunchKernelGGL(k_envr2, dim3((unsigned)grid_for(w2, 256)), dim3(256), 0, stream_, sites, off_dev_.p,
                       caps_dev_.p, dims, j, envT_.p, rout, ldr);
    TQ_LAUNCH_CHECK();
}

void Tebd::env_left(const double2* lin, int j, bool z, double2* lout) {
    const int ldl = chi_max();
    const long long w1 = (long long)cap_[j] * 2 * cap_[j + 1];
    hipLaunchKernelGGL(k_envl1, dim3((unsigned)grid_for(w1, 256)), dim3(256), 0, stream_, sites_.p, off_dev_.p,
                       caps_dev_.p, dims_.p, j, lin, ldl, envT_.p);
    TQ_LAUNCH_CHECK();
    const long long w2 = (long long)cap_[j + 1] * cap_[j + 1];
    hipLaunchKernelGGL(k_envl2, dim3((unsigned)grid_for(w2, 256)), dim3(256), 0, stream_, sites_.p, off_dev_.p,
                       caps_dev_.p, dims_.p, j, z ? 1 : 0, envT_.p, lout, ldl);
    TQ_LAUNCH_CHECK();
}

// mps.cpp:355 -> single_site_entropy(0): right environment over sites n-1..1.
void Tebd::entropy_sweep(int rec) { entropy_of(sites_.p, dims_.p, rec); }

void Tebd::entropy_of(const double2* sites, const int* dims, int rec) {
    ensure_trace(rec + 1);
    entropy_into(sites, dims, trace_.p + rec);
}

double Tebd::entropy_now() {
    DevBuf<double> out(1);
    entropy_into(sites_.p, dims_.p, out.p);
    double s = 0.0;
    TQ_HIP(hipMemcpyAsync(&s, out.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    return s;
}

void Tebd::entropy_into(const double2* sites, const int* dims, double* dst) {
    if (n_ < 2) {
        const double zero = 0.0;
        TQ_HIP(hipMemcpyAsync(dst, &zero, sizeof(double), hipMemcpyHostToDevice, stream_));
        TQ_HIP(hipStreamSynchronize(stream_));
        return;
    }
    double2* cur = envA_.p;
    double2* nxt = envB_.p;
    hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, stream_, cur);
    TQ_LAUNCH_CHECK();
    for (int j = n_ - 1; j >= 1; --j) {
        env_right(cur, j, nxt, sites, dims);
        std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_rho0, dim3(1), dim3(64), 0, stream_, sites, off_dev_.p, caps_dev_.p, dims, cur,
                       chi_max(), dst, 0, capturing_ ? sp_.par : nullptr);
    TQ_LAUNCH_CHECK();
}

void Tebd::sync() {
    TQ_HIP(hipStreamSynchronize(stream_));
    for (std::size_t q = 0; q < ev_next_; ++q) {  // fold recorded SVD launch times
        float ms = 0.f;
        TQ_HIP(hipEventElapsedTime(&ms, events_[q].first, events_[q].second));
        timed_ms_ += ms;
        ++timed_n_;
    }
    ev_next_ = 0;
    if (ex_armed_) {  // a group barrier of the last executor launch that gave up
        ex_armed_ = false;
        int err = 0;
        TQ_HIP(hipMemcpy(&err, ex_bar_.p, sizeof(int), hipMemcpyDeviceToHost));
        if (err & 2)
            throw Error(-3, "lockstep executor: workgroups were not placed one group per XCD; the launch did nothing");
        if (err) throw Error(-3, "lockstep executor: a group barrier wait timed out; results are invalid");
    }
}

std::vector<double> Tebd::exec_profile() {
    std::vector<double> out(kProfSlots, 0.0);
    if (!ex_prof_.p) return out;
    std::vector<unsigned long long> h(kProfSlots);
    TQ_HIP(hipMemcpyAsync(h.data(), ex_prof_.p, kProfSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          stream_));
    sync();
    int dev = 0, khz = 0;
    TQ_HIP(hipGetDevice(&dev));
    TQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const double ms_per_tick = khz > 0 ? 1.0 / khz : 0.0;
    for (int q = 0; q < kProfSlots; ++q)
        out[q] = (q == 7 || q == 9 || q == 11 || (q >= kProfHist && q < kProfRound)) ? (double)h[q]
                                                                                  : (double)h[q] * ms_per_tick;
    return out;
}

void Tebd::set_timing(bool on) {
    sync();
    timing_ = on;
    timed_ms_ = 0.0;
    timed_n_ = 0;
}

std::vector<double> Tebd::entropy(int count) {
    std::vector<double> out(count, 0.0);
    if (count > 0 && trace_cap_ >= count)
        TQ_HIP(hipMemcpyAsync(out.data(), trace_.p, count * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    return out;
}

SvdStats Tebd::svd_stats() {
    std::vector<unsigned long long> h(4 * (std::size_t)kStatSlots);
    TQ_HIP(hipMemcpyAsync(h.data(), stats_.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
    sync();
    unsigned long long tot[4] = {0, 0, 0, 0};
    for (int b = 0; b < kStatSlots; ++b)
        for (int k = 0; k < 4; ++k) tot[k] += h[4 * (std::size_t)b + k];
    SvdStats st;
    st.calls = (long long)tot[0];
    st.sweeps = (long long)tot[1];
    st.visited_rows = (double)tot[2];
    st.rotated_rows = (double)tot[3];
    st.svd_ms = timed_ms_;
    st.timed_launches = timed_n_;
    st.supersteps = exec_supersteps_;
    st.exec_updates = exec_updates_;
    return st;
}

double Tebd::max_trunc() {
    std::vector<double> t(kStatSlots);
    TQ_HIP(hipMemcpyAsync(t.data(), trunc_.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    double m = 0.0;
    for (double v : t) m = std::max(m, v);
    return m;
}

std::vector<int> Tebd::bonds() {
    std::vector<int> d(n_ + 1);
    TQ_HIP(hipMemcpyAsync(d.data(), dims_.p, (n_ + 1) * sizeof(int), hipMemcpyDeviceToHost, stream_));
    sync();
    return d;
}

// mps.cpp:358-380: <H> from normalised <Z_i Z_j> and <Z_i>, spins = sign<Z_i>
// (0 -> +1), readout energy in edge-then-field order, prob = |<x|psi>|^2 / norm.
Tebd::Readout Tebd::readout() {
    const int ld = chi_max();
    const std::size_t blk = (std::size_t)ld * ld;
    DevBuf<double2> L((std::size_t)(n_ + 1) * blk), Rr((std::size_t)(n_ + 1) * blk), X(blk), Y(blk);
    hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, stream_, L.p);
    hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, stream_, Rr.p + (std::size_t)n_ * blk);
    TQ_LAUNCH_CHECK();
    for (int i = 0; i < n_; ++i) env_left(L.p + (std::size_t)i * blk, i, false, L.p + (std::size_t)(i + 1) * blk);
    for (int j = n_ - 1; j >= 0; --j) env_right(Rr.p + (std::size_t)(j + 1) * blk, j, Rr.p + (std::size_t)j * blk);
    const int m = (int)model_.ew.size();
    // slots: 0 norm, 1..n <Z_i>, n+1..n+m <ZZ_e>, n+m+1 amplitude
    DevBuf<double2> res(n_ + m + 2);
    hipLaunchKernelGGL(k_close, dim3(1), dim3(256), 0, stream_, L.p + (std::size_t)n_ * blk, Rr.p + (std::size_t)n_ * blk,
                       ld, dims_.p, n_, res.p, 0);
    for (int i = 0; i < n_; ++i) {
        env_left(L.p + (std::size_t)i * blk, i, true, X.p);
        hipLaunchKernelGGL(k_close, dim3(1), dim3(256), 0, stream_, X.p, Rr.p + (std::size_t)(i + 1) * blk, ld,
                           dims_.p, i + 1, res.p, 1 + i);
    }
    for (int e = 0; e < m; ++e) {
        int i = model_.ei[e], j = model_.ej[e];
        if (i > j) std::swap(i, j);
        double2* a = X.p;
        double2* b = Y.p;
        env_left(L.p + (std::size_t)i * blk, i, true, a);
        for (int k = i + 1; k <= j; ++k) {
            env_left(a, k, k == j, b);
            std::swap(a, b);
        }
        // i == j: a single Z at i, as the upstream z-string membership test does
        hipLaunchKernelGGL(k_close, dim3(1), dim3(256), 0, stream_, a, Rr.p + (std::size_t)(j + 1) * blk, ld, dims_.p,
                           j + 1, res.p, 1 + n_ + e);
    }
    TQ_LAUNCH_CHECK();
    std::vector<double2> hres(n_ + m + 2);
    TQ_HIP(hipMemcpyAsync(hres.data(), res.p, (n_ + m + 1) * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    sync();
    const double nrm = hres[0].x;
    Readout out;
    std::vector<double> z(n_), zz(m);
    for (int i = 0; i < n_; ++i) z[i] = nrm <= 0 ? 0.0 : hres[1 + i].x / nrm;
    for (int e = 0; e < m; ++e) zz[e] = nrm <= 0 ? 0.0 : hres[1 + n_ + e].x / nrm;
    double eexp = 0.0;
    for (int e = 0; e < m; ++e) eexp -= model_.ew[e] * zz[e];
    for (int i = 0; i < n_; ++i) eexp -= model_.h[i] * z[i];
    out.expected_energy = eexp;
    out.spins.resize(n_);
    for (int i = 0; i < n_; ++i) out.spins[i] = z[i] >= 0.0 ? 1 : -1;
    double en = 0.0;
    for (int e = 0; e < m; ++e) en -= model_.ew[e] * out.spins[model_.ei[e]] * out.spins[model_.ej[e]];
    for (int i = 0; i < n_; ++i) en -= model_.h[i] * out.spins[i];
    out.energy = en;
    DevBuf<int8_t> dsp(n_);
    TQ_HIP(hipMemcpyAsync(dsp.p, out.spins.data(), n_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_amplitude, dim3(1), dim3(256), 0, stream_, sites_.p, off_dev_.p, caps_dev_.p, dims_.p, dsp.p,
                       n_, res.p + n_ + m + 1);
    TQ_LAUNCH_CHECK();
    double2 amp;
    TQ_HIP(hipMemcpyAsync(&amp, res.p + n_ + m + 1, sizeof(double2), hipMemcpyDeviceToHost, stream_));
    sync();
    out.prob = nrm > 0 ? (amp.x * amp.x + amp.y * amp.y) / nrm : 0.0;
    out.z = std::move(z);
    return out;
}

void Tebd::set_bias(const double* bias) {
    sync();  // launches already queued may still read the old field
    if (!bias) {
        model_.bias.clear();
        bias_.release();
        return;
    }
    model_.bias.assign(bias, bias + n_);
    if (!bias_.p) bias_.alloc(n_);
    TQ_HIP(hipMemcpy(bias_.p, model_.bias.data(), n_ * sizeof(double), hipMemcpyHostToDevice));
}

// theta of sites (p, p+1) under the identity gate, as the SVD sees it: R x C
// = 2Dl x 2Dr row-major complex (interleaved re/im), R and C in rows/cols.
std::vector<double> Tebd::theta(int p, int& rows, int& cols) {
    if (p < 0 || p + 1 >= n_) throw Error(-1, "theta site out of range");
    Gate4 eye;
    for (auto& x : eye.g) x = make_double2(0, 0);
    for (int k = 0; k < 4; ++k) eye.g[k * 5] = make_double2(1, 0);
    const std::vector<int> d = bonds();
    const int Dl = d[p], Dr = d[p + 2];
    rows = 2 * Dl;
    cols = 2 * Dr;
    const long long ttiles = theta_tiles_for(cap_[p], cap_[p + 2]);
    hipLaunchKernelGGL(k_theta, dim3((unsigned)((ttiles + 3) / 4)), dim3(256), 0, stream_, sites_.p, off_dev_.p,
                       caps_dev_.p, dims_.p, p, eye, mat_.p, work_.p, ld_, StepParams{}, -1);
    TQ_LAUNCH_CHECK();
    std::vector<double2> m((std::size_t)ld_ * ld_);
    TQ_HIP(hipMemcpyAsync(m.data(), mat_.p, m.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    sync();
    const bool tall = cols <= rows;  // M = theta column-major, else theta^dagger column-major
    std::vector<double> out(2 * (std::size_t)rows * cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const double2 v = tall ? m[(std::size_t)c * ld_ + r] : m[(std::size_t)r * ld_ + c];
            out[2 * ((std::size_t)r * cols + c)] = v.x;
            out[2 * ((std::size_t)r * cols + c) + 1] = tall ? v.y : -v.y;
        }
    return out;
}

std::vector<double> Tebd::dense() {
    if (n_ > 24) throw Error(-1, "dense state only for n <= 24");
    std::vector<double2> host(off_[n_]);
    TQ_HIP(hipMemcpyAsync(host.data(), sites_.p, host.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    std::vector<int> d = bonds();
    std::vector<std::complex<double>> cur(1, {1.0, 0.0});
    std::size_t rows = 1;
    int bond = 1;
    for (int i = 0; i < n_; ++i) {
        const int Dr = d[i + 1], st = cap_[i + 1];
        std::vector<std::complex<double>> nxt(rows * 2 * Dr);
        for (std::size_t x = 0; x < rows; ++x)
            for (int s = 0; s < 2; ++s)
                for (int b = 0; b < Dr; ++b) {
                    std::complex<double> acc(0, 0);
                    for (int a = 0; a < bond; ++a) {
                        const double2 t = host[off_[i] + ((std::size_t)a * 2 + s) * st + b];
                        acc += cur[x * bond + a] * std::complex<double>(t.x, t.y);
                    }
                    nxt[(x + ((std::size_t)s << i)) * Dr + b] = acc;
                }
        cur.swap(nxt);
        rows *= 2;
        bond = Dr;
    }
    std::vector<double> out(2 * cur.size());
    for (std::size_t x = 0; x < cur.size(); ++x) {
        out[2 * x] = cur[x].real();
        out[2 * x + 1] = cur[x].imag();
    }
    return out;
}

}  // namespace tq
