// Exact state-vector anneal on gfx950 (upstream rust/src/quantum.rs:63-112).
//
// Layout: psi[x] as interleaved complex128 (double2), x = sum_i bit_i 2^i, bit
// set <-> spin +1. The problem diagonal is never stored: E(x) is recomputed from
// the index bits in the upstream summation order (edges in list order, then
// fields, ising.rs:23-32), so it is bit-identical to quantum.rs:46-56.
//
// Per Trotter step k (quantum.rs:80-87):
//   phase(b dt/2) -> X rotations on every qubit (angle a dt) -> phase(b dt/2)
//   -> qubit-0 entropy and <H_problem>.
// n <= 13: one workgroup keeps the whole state in LDS and runs all steps.
// n >= 14: each step streams the state through HBM in passes over 2^12-amplitude
//   tiles. A tile is 2^c contiguous amplitudes (c >= 3: >= 128 B runs) times a
//   window of 12-c higher bits; the first pass rotates bits 0..11 of contiguous
//   4096-blocks, later passes rotate their window. Inside a tile each of 256
//   threads holds 16 amplitudes; rotations on register bits need no exchange,
//   and LDS exchanges cycle local bits 0-3, 4-7, 8-11 into registers, so each
//   amplitude sees its rotations in ascending qubit order as upstream does.
// Every multiply and add is rounded separately (no FMA contraction) to follow
// the upstream operation order; rotation cos/sin are computed on the host with
// the same libm calls as upstream, the phase cos/sin per amplitude on the device.
#include <cmath>

#include "common.hpp"
#include "sv.hpp"

#pragma clang fp contract(off)

namespace tq {

namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kThreads = 256;
constexpr int kRegs = kTile / kThreads;  // 16 amplitudes per thread
constexpr int kResidentMaxN = 13;
constexpr int kObs = 6;  // r00, r11, re r01, im r01, sum p, sum p*E

struct Model {
    int n, m;
    const int* ei;
    const int* ej;
    const double* ew;
    const double* h;
};

// quantum.rs:46-56 via ising.rs:23-32, bit set => +1.
__device__ inline double energy_of(const Model& md, unsigned long long x) {
    double e = 0.0;
    for (int k = 0; k < md.m; ++k) {
        const double si = ((x >> md.ei[k]) & 1ull) ? 1.0 : -1.0;
        const double sj = ((x >> md.ej[k]) & 1ull) ? 1.0 : -1.0;
        e -= md.ew[k] * si * sj;
    }
    for (int i = 0; i < md.n; ++i) {
        const double si = ((x >> i) & 1ull) ? 1.0 : -1.0;
        e -= md.h[i] * si;
    }
    return e;
}

// quantum.rs:178-187
__device__ inline double2 apply_phase(double2 v, double scale, double e) {
    const double phi = scale * e;
    double s, c;
    sincos(phi, &s, &c);
    return make_double2(v.x * c + v.y * s, v.y * c - v.x * s);
}

// quantum.rs:196-203: (a, b) = (x, x|bit)
__device__ inline void rotate_pair(double2& a, double2& b, double c, double s) {
    const double ra = a.x, ia = a.y, rb = b.x, ib = b.y;
    a = make_double2(c * ra - s * ib, c * ia + s * rb);
    b = make_double2(c * rb - s * ia, c * ib + s * ra);
}

__device__ inline int sidx(int j) { return j + (j >> 4); }  // LDS padding: 1 slot per 16

// Rotate the register bits selected by `mask4` (bit i of mask4 <-> register bit i).
__device__ inline void rotate_regs(double2 (&v)[kRegs], unsigned mask4, double c, double s) {
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        if (!((mask4 >> bit) & 1u)) continue;
#pragma unroll
        for (int r = 0; r < kRegs; ++r)
            if (!(r & (1 << bit))) rotate_pair(v[r], v[r | (1 << bit)], c, s);
    }
}

// Local index layouts (t = thread, r = register slot).
__device__ inline int lay0(int t, int r) { return t + (r << 8); }                          // regs <-> bits 8..11
__device__ inline int lay1(int t, int r) { return (t & 15) | (r << 4) | ((t >> 4) << 8); }  // regs <-> bits 4..7
__device__ inline int lay2(int t, int r) { return r | (t << 4); }                          // regs <-> bits 0..3

// ---------------------------------------------------------------------------
// n <= 13: whole anneal in one workgroup, state resident in LDS.
__global__ void __launch_bounds__(1024) k_sv_resident(Model md, const double* sa, const double* sb,
                                                      const double* rot_c, const double* rot_s, double dt,
                                                      int steps, double amp0, double* ent_out,
                                                      double* en_out, double2* state_out) {
    __shared__ double2 st[1 << kResidentMaxN];
    __shared__ double red[16][kObs];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int n = md.n;
    const int dim = 1 << n;
    for (int x = tid; x < dim; x += nth) st[x] = make_double2(amp0, 0.0);
    __syncthreads();
    for (int k = 0; k < steps; ++k) {
        const double scale = sb[k] * dt * 0.5;
        const double c = rot_c[k], s = rot_s[k];
        for (int x = tid; x < dim; x += nth) st[x] = apply_phase(st[x], scale, energy_of(md, (unsigned)x));
        __syncthreads();
        for (int q = 0; q < n; ++q) {
            for (int p = tid; p < (dim >> 1); p += nth) {
                const int lowm = (1 << q) - 1;
                const int x = ((p & ~lowm) << 1) | (p & lowm);
                double2 a = st[x], b = st[x | (1 << q)];
                rotate_pair(a, b, c, s);
                st[x] = a;
                st[x | (1 << q)] = b;
            }
            __syncthreads();
        }
        double acc[kObs] = {0, 0, 0, 0, 0, 0};
        for (int x = tid; x < dim; x += nth) {
            const double e = energy_of(md, (unsigned)x);
            const double2 v = apply_phase(st[x], scale, e);
            st[x] = v;
            const double p = v.x * v.x + v.y * v.y;
            acc[4] += p;
            acc[5] += p * e;
        }
        __syncthreads();
        if (n >= 2)
            for (int x = 2 * tid; x < dim; x += 2 * nth) {
                const double2 u = st[x], w = st[x + 1];
                acc[0] += u.x * u.x + u.y * u.y;
                acc[1] += w.x * w.x + w.y * w.y;
                acc[2] += u.x * w.x + u.y * w.y;
                acc[3] += u.y * w.x - u.x * w.y;
            }
#pragma unroll
        for (int q = 0; q < kObs; ++q) acc[q] = wave_sum(acc[q]);
        if ((tid & 63) == 0)
            for (int q = 0; q < kObs; ++q) red[tid >> 6][q] = acc[q];
        __syncthreads();
        if (tid == 0) {
            double tot[kObs] = {0, 0, 0, 0, 0, 0};
            for (int w = 0; w < (nth >> 6); ++w)
                for (int q = 0; q < kObs; ++q) tot[q] += red[w][q];
            ent_out[k] = n < 2 ? 0.0 : sv_entropy_from(tot[0], tot[1], tot[2], tot[3]);
            en_out[k] = tot[5] / tot[4];
        }
        __syncthreads();
    }
    for (int x = tid; x < dim; x += nth) state_out[x] = st[x];
}

// ---------------------------------------------------------------------------
// n >= 14: one streaming pass. `c` contiguous low bits, window starting at `lo`;
// local bits set in `rotmask` are rotated. PHASE_IN applies phase(scale_in)
// after the load; PHASE_OUT applies phase(scale_out) after the rotations and
// accumulates the step observables into partials[block][6].
template <bool PHASE_IN, bool PHASE_OUT>
__global__ void __launch_bounds__(kThreads) k_sv_pass(double2* __restrict__ psi, Model md, int c, int lo,
                                                      unsigned rotmask, double rc, double rs,
                                                      double scale_in, double scale_out,
                                                      long long ntiles, double* __restrict__ partials) {
    __shared__ double2 sh[kTile + kTile / 16];
    __shared__ double red[kThreads / 64][kObs];
    const int t = threadIdx.x;
    const int lo_span = lo - c;  // tile-id bits placed in [c, lo)
    const unsigned long long cmask = (1ull << c) - 1ull;
    double acc[kObs] = {0, 0, 0, 0, 0, 0};

    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        unsigned long long base;
        if (lo_span <= 0) {
            base = (unsigned long long)tile << kTileBits;
        } else {
            const unsigned long long tl = (unsigned long long)tile & ((1ull << lo_span) - 1ull);
            const unsigned long long th = (unsigned long long)tile >> lo_span;
            base = (tl << c) | (th << (lo + (kTileBits - c)));
        }
        auto gidx = [&](int j) -> unsigned long long {
            return base | ((unsigned long long)j & cmask) | (((unsigned long long)j >> c) << lo);
        };
        auto gidx_low = [&](int j) -> unsigned long long { return base | (unsigned long long)j; };

        double2 v[kRegs];
#pragma unroll
        for (int r = 0; r < kRegs; ++r) {
            const int j = lay0(t, r);
            v[r] = psi[lo_span < 0 ? gidx_low(j) : gidx(j)];
        }
        if (PHASE_IN) {
#pragma unroll
            for (int r = 0; r < kRegs; ++r) {
                const int j = lay0(t, r);
                const unsigned long long x = lo_span < 0 ? gidx_low(j) : gidx(j);
                v[r] = apply_phase(v[r], scale_in, energy_of(md, x));
            }
        }
        // ascending qubit order as quantum.rs:193: local bits 0-3, 4-7, 8-11
#pragma unroll
        for (int r = 0; r < kRegs; ++r) sh[sidx(lay0(t, r))] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay2(t, r))];
        __syncthreads();
        rotate_regs(v, rotmask & 15u, rc, rs);
#pragma unroll
        for (int r = 0; r < kRegs; ++r) sh[sidx(lay2(t, r))] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay1(t, r))];
        __syncthreads();
        rotate_regs(v, (rotmask >> 4) & 15u, rc, rs);
#pragma unroll
        for (int r = 0; r < kRegs; ++r) sh[sidx(lay1(t, r))] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay0(t, r))];
        __syncthreads();
        rotate_regs(v, (rotmask >> 8) & 15u, rc, rs);
        if (PHASE_OUT) {
            const bool even = (t & 1) == 0;  // local bit 0 == global bit 0 == thread bit 0
#pragma unroll
            for (int r = 0; r < kRegs; ++r) {
                const unsigned long long x = lo_span < 0 ? gidx_low(lay0(t, r)) : gidx(lay0(t, r));
                const double e = energy_of(md, x);
                v[r] = apply_phase(v[r], scale_out, e);
                const double p = v[r].x * v[r].x + v[r].y * v[r].y;
                acc[4] += p;
                acc[5] += p * e;
                const double2 w = make_double2(__shfl_xor(v[r].x, 1, 64), __shfl_xor(v[r].y, 1, 64));
                if (even) {
                    const double2 u = v[r];
                    acc[0] += u.x * u.x + u.y * u.y;
                    acc[1] += w.x * w.x + w.y * w.y;
                    acc[2] += u.x * w.x + u.y * w.y;
                    acc[3] += u.y * w.x - u.x * w.y;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kRegs; ++r) {
            const int j = lay0(t, r);
            psi[lo_span < 0 ? gidx_low(j) : gidx(j)] = v[r];
        }
    }
    if (PHASE_OUT) {
#pragma unroll
        for (int q = 0; q < kObs; ++q) acc[q] = wave_sum(acc[q]);
        if ((t & 63) == 0)
            for (int q = 0; q < kObs; ++q) red[t >> 6][q] = acc[q];
        __syncthreads();
        if (t == 0)
            for (int q = 0; q < kObs; ++q) {
                double s = 0.0;
                for (int w = 0; w < kThreads / 64; ++w) s += red[w][q];
                partials[(long long)blockIdx.x * kObs + q] = s;
            }
    }
}

// Fixed-order sum of per-block partials -> entropy and <H_problem> of step k.
__global__ void __launch_bounds__(256) k_sv_finish_step(const double* __restrict__ partials, int nblocks, int n,
                                                        int k, double* ent_out, double* en_out) {
    __shared__ double red[4][kObs];
    double acc[kObs] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int q = 0; q < kObs; ++q) acc[q] += partials[(long long)b * kObs + q];
    for (int q = 0; q < kObs; ++q) acc[q] = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < kObs; ++q) red[threadIdx.x >> 6][q] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[kObs] = {0, 0, 0, 0, 0, 0};
        for (int w = 0; w < 4; ++w)
            for (int q = 0; q < kObs; ++q) tot[q] += red[w][q];
        ent_out[k] = n < 2 ? 0.0 : sv_entropy_from(tot[0], tot[1], tot[2], tot[3]);
        en_out[k] = tot[5] / tot[4];
    }
}

__global__ void k_sv_fill(double2* psi, long long dim, double amp0) {
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < dim; x += (long long)gridDim.x * blockDim.x)
        psi[x] = make_double2(amp0, 0.0);
}

// Final measurement statistics (quantum.rs:89-102): total, sum p*E, and the
// argmax with the lowest index among equal maxima.
struct Best {
    double p;
    long long x;
};
__device__ inline Best better(Best a, Best b) {
    if (b.p > a.p || (b.p == a.p && b.x < a.x)) return b;
    return a;
}

__global__ void __launch_bounds__(256) k_sv_measure(const double2* __restrict__ psi, Model md, long long dim,
                                                    double* __restrict__ part_sum, long long* __restrict__ part_idx) {
    __shared__ double rs[4][3];
    __shared__ long long ri[4];
    double tot = 0.0, ex = 0.0;
    Best best{-1.0, 0};
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < dim;
         x += (long long)gridDim.x * blockDim.x) {
        const double2 v = psi[x];
        const double p = v.x * v.x + v.y * v.y;
        tot += p;
        ex += p * energy_of(md, (unsigned long long)x);
        best = better(best, Best{p, x});
    }
    tot = wave_sum(tot);
    ex = wave_sum(ex);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Best o{__shfl_xor(best.p, off, 64), __shfl_xor(best.x, off, 64)};
        best = better(best, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        rs[w][0] = tot;
        rs[w][1] = ex;
        rs[w][2] = best.p;
        ri[w] = best.x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t2 = 0.0, e2 = 0.0;
        Best b2{-1.0, 0};
        for (int q = 0; q < 4; ++q) {
            t2 += rs[q][0];
            e2 += rs[q][1];
            b2 = better(b2, Best{rs[q][2], ri[q]});
        }
        part_sum[blockIdx.x * 3 + 0] = t2;
        part_sum[blockIdx.x * 3 + 1] = e2;
        part_sum[blockIdx.x * 3 + 2] = b2.p;
        part_idx[blockIdx.x] = b2.x;
    }
}

}  // namespace

// quantum.rs:138-175 closed form (shared with the host-side tests).
__host__ __device__ double sv_entropy_from(double r00, double r11, double r01r, double r01i) {
    const double total = r00 + r11;
    if (total <= 0.0) return 0.0;
    r00 /= total;
    r11 /= total;
    const double off2 = (r01r * r01r + r01i * r01i) / (total * total);
    const double mean = (r00 + r11) / 2.0;
    const double half = (r00 - r11) / 2.0;
    const double rad = sqrt(half * half + off2);
    double lp = mean + rad, lm = mean - rad;
    lp = lp < 0.0 ? 0.0 : (lp > 1.0 ? 1.0 : lp);
    lm = lm < 0.0 ? 0.0 : (lm > 1.0 ? 1.0 : lm);
    double s = 0.0;
    if (lp > 1e-12) s -= lp * log2(lp);
    if (lm > 1e-12) s -= lm * log2(lm);
    return s;
}

SvOutcome sv_anneal_device(const SvProblem& pb, bool keep_state) {
    const int n = pb.n;
    if (n < 1 || n > 40) throw Error(-1, "state-vector anneal needs 1 <= n <= 40");
    if (pb.steps < 1) throw Error(-1, "state-vector anneal needs steps >= 1");
    const long long dim = 1ll << n;
    const int m = (int)pb.ew.size();
    SvOutcome out;
    out.entropy.assign(pb.steps, 0.0);
    out.energy_trace.assign(pb.steps, 0.0);

    // model upload
    DevBuf<int> d_ei(std::max(m, 1)), d_ej(std::max(m, 1));
    DevBuf<double> d_ew(std::max(m, 1)), d_h(n);
    if (m) {
        TQ_HIP(hipMemcpy(d_ei.p, pb.ei.data(), m * sizeof(int), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_ej.p, pb.ej.data(), m * sizeof(int), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_ew.p, pb.ew.data(), m * sizeof(double), hipMemcpyHostToDevice));
    }
    TQ_HIP(hipMemcpy(d_h.p, pb.h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    Model md{n, m, d_ei.p, d_ej.p, d_ew.p, d_h.p};

    // per-step rotation cos/sin from the host libm, as quantum.rs:191
    std::vector<double> rc(pb.steps), rsn(pb.steps);
    for (int k = 0; k < pb.steps; ++k) {
        const double th = pb.a[k] * pb.dt;
        rc[k] = std::cos(th);
        rsn[k] = std::sin(th);
    }
    const double amp0 = 1.0 / std::sqrt((double)dim);  // quantum.rs:72

    DevBuf<double2> psi(dim);
    DevBuf<double> d_ent(pb.steps), d_en(pb.steps);
    hipStream_t st = nullptr;

    if (n <= kResidentMaxN) {
        DevBuf<double> d_a(pb.steps), d_b(pb.steps), d_c(pb.steps), d_s(pb.steps);
        TQ_HIP(hipMemcpy(d_b.p, pb.b.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_c.p, rc.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_s.p, rsn.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_sv_resident, dim3(1), dim3(1024), 0, st, md, d_a.p, d_b.p, d_c.p, d_s.p, pb.dt,
                           pb.steps, amp0, d_ent.p, d_en.p, psi.p);
        TQ_LAUNCH_CHECK();
    } else {
        const long long ntiles = dim >> kTileBits;
        const int nblocks = (int)std::min<long long>(ntiles, 4096);
        DevBuf<double> partials((std::size_t)nblocks * kObs);
        hipLaunchKernelGGL(k_sv_fill, dim3(2048), dim3(256), 0, st, psi.p, dim, amp0);
        TQ_LAUNCH_CHECK();
        // pass plan: low pass over bits [0,12), then windows of <= 9 bits
        struct Pass {
            int c, lo;
            unsigned mask;
        };
        std::vector<Pass> plan;
        plan.push_back({kTileBits, 0, (1u << kTileBits) - 1u});
        for (int lo = kTileBits; lo < n; lo += 9) {
            const int w = std::min(9, n - lo);
            const int c = kTileBits - w;
            plan.push_back({c, lo, ((1u << w) - 1u) << c});
        }
        for (int k = 0; k < pb.steps; ++k) {
            const double scale = pb.b[k] * pb.dt * 0.5;
            for (std::size_t p = 0; p < plan.size(); ++p) {
                const Pass& ps = plan[p];
                const bool first = p == 0, last = p + 1 == plan.size();
                const int lo = first ? -1 : ps.lo;  // lo < c selects the contiguous-block addressing
                const int c = first ? 0 : ps.c;
                if (first)
                    hipLaunchKernelGGL((k_sv_pass<true, false>), dim3(nblocks), dim3(kThreads), 0, st, psi.p, md,
                                       c, lo, ps.mask, rc[k], rsn[k], scale, scale, ntiles, partials.p);
                else if (last)
                    hipLaunchKernelGGL((k_sv_pass<false, true>), dim3(nblocks), dim3(kThreads), 0, st, psi.p, md,
                                       c, lo, ps.mask, rc[k], rsn[k], scale, scale, ntiles, partials.p);
                else
                    hipLaunchKernelGGL((k_sv_pass<false, false>), dim3(nblocks), dim3(kThreads), 0, st, psi.p, md,
                                       c, lo, ps.mask, rc[k], rsn[k], scale, scale, ntiles, partials.p);
                TQ_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_sv_finish_step, dim3(1), dim3(256), 0, st, partials.p, nblocks, n, k, d_ent.p,
                               d_en.p);
            TQ_LAUNCH_CHECK();
        }
    }

    const int mblocks = (int)std::min<long long>((dim + 255) / 256, 2048);
    DevBuf<double> ps(mblocks * 3);
    DevBuf<long long> pi(mblocks);
    hipLaunchKernelGGL(k_sv_measure, dim3(mblocks), dim3(256), 0, st, psi.p, md, dim, ps.p, pi.p);
    TQ_LAUNCH_CHECK();
    TQ_HIP(hipStreamSynchronize(st));
    TQ_HIP(hipMemcpy(out.entropy.data(), d_ent.p, pb.steps * sizeof(double), hipMemcpyDeviceToHost));
    TQ_HIP(hipMemcpy(out.energy_trace.data(), d_en.p, pb.steps * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> hs(mblocks * 3);
    std::vector<long long> hi(mblocks);
    TQ_HIP(hipMemcpy(hs.data(), ps.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
    TQ_HIP(hipMemcpy(hi.data(), pi.p, hi.size() * sizeof(long long), hipMemcpyDeviceToHost));
    double total = 0.0, expected = 0.0, best_p = -1.0;
    long long best = 0;
    for (int b = 0; b < mblocks; ++b) {
        total += hs[b * 3 + 0];
        expected += hs[b * 3 + 1];
        if (hs[b * 3 + 2] > best_p || (hs[b * 3 + 2] == best_p && hi[b] < best)) {
            best_p = hs[b * 3 + 2];
            best = hi[b];
        }
    }
    out.spins.resize(n);
    for (int i = 0; i < n; ++i) out.spins[i] = ((best >> i) & 1) ? 1 : -1;
    // energy of the argmax in the upstream order (quantum.rs:106)
    double e = 0.0;
    for (int k = 0; k < m; ++k) e -= pb.ew[k] * (double)out.spins[pb.ei[k]] * (double)out.spins[pb.ej[k]];
    for (int i = 0; i < n; ++i) e -= pb.h[i] * (double)out.spins[i];
    out.energy = e;
    out.expected_energy = expected / total;
    out.success_prob = best_p / total;
    out.best_index = best;
    if (keep_state) {
        out.state.resize(2 * (std::size_t)dim);
        TQ_HIP(hipMemcpy(out.state.data(), psi.p, dim * sizeof(double2), hipMemcpyDeviceToHost));
    }
    return out;
}

}  // namespace tq
