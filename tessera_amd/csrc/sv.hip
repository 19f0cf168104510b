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
// Rotation cos/sin are computed on the host with the same libm calls as upstream;
// phases come from per-tile tables (below), so results agree with upstream to
// rounding (not bitwise).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "sv.hpp"


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

// Energy terms of one pass, classified by where their bits sit in a tile:
// L = local bits 0-5, H = local bits 6-11, B = bits fixed by the tile (base).
// Per tile E(base|j) = C + T1[j & 63] + T2[j >> 6] - sum_{LH} w s s, where T1 holds
// the LL, BL edges and L fields, T2 the HH, BH edges and H fields, C the BB
// edges and B fields. Edge terms are (a, b, w); field terms (a, -, h).
struct Term {
    int a, b;
    double w;
};
enum { kLL, kHH, kBL, kBH, kBB, kLH, kFL, kFH, kFB, kKinds };
struct PassTerms {
    const Term* t;
    int off[kKinds + 1];
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
__global__ void __launch_bounds__(1024) k_sv_resident(Model md, const double* sb,
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
// n >= 14: streaming passes over 2^12-amplitude tiles.
//
// The X rotations of different qubits commute, so a step may rotate its bit
// groups in any order; steps alternate ascending and descending group order so
// that the last group of step k is the first group of step k+1. That shared
// group is done in one pass: rotate(k), phase(k), observables(k), phase(k+1),
// rotate(k+1). With g groups a step costs g-1 passes (2 at n = 30).
//
// Phases come from per-tile tables: bits of a tile split into L (local 0-5),
// H (local 6-11) and B (fixed by the tile); E(base|j) = C + T1[j & 63] +
// T2[j >> 6] - sum_{LH edges} w s s and exp(-i scale E) is the product of the
// matching table phasors, so each amplitude costs a few complex products and no
// transcendental.
constexpr int kMaxTerms = 256;
constexpr int kMaxCrossT = 64;
// Register phase block: up to this many L-H couplings (their 2^k sign patterns
// are tabulated once per pass); more take the LDS phase block.
constexpr int kFastCross = 6;

struct PassArgs {
    int c, lo;         // tile: 2^c contiguous amplitudes, window at lo (lo < 0: contiguous 4096 block)
    unsigned mask;     // local bits of the rotation group
    int rot_a, rot_b;  // rotate before / after the phase block
    double ca, sa, cb, sb;
    int phase_out, phase_in;
    double scale_out, scale_in;
    int best;          // track the argmax (last step)
    int fast;          // register phase block (L-H couplings <= kFastCross)
    int nx;            // L-H couplings: local bits xa[q] (L), xb[q] (H), weight xw[q]
    int xa[kFastCross], xb[kFastCross];
    double xw[kFastCross];
    int x0mask;        // couplings that involve local bit 0 (flip with it)
};

__device__ inline double bit_spin(unsigned long long x, int b) { return ((x >> b) & 1ull) ? 1.0 : -1.0; }

struct Best {
    double p;
    long long x;
};
__device__ inline Best better(Best a, Best b) {
    if (b.p > a.p || (b.p == a.p && b.x < a.x)) return b;
    return a;
}

// Amplitudes held in registers: a plain two-double struct (arrays of HIP's
// union-based double2 are kept from register promotion in some loop forms).
struct alignas(16) c2 {
    double x, y;
};
__device__ inline c2 ldc(const double2* p) {
    const double2 q = *p;
    return c2{q.x, q.y};
}
__device__ inline void stc(double2* p, c2 v) { *p = make_double2(v.x, v.y); }
__device__ inline c2 cm(c2 a, c2 b) { return c2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline c2 cj(c2 a) { return c2{a.x, -a.y}; }
__device__ inline c2 c2of(double2 a) { return c2{a.x, a.y}; }

// The value of lane t ^ 1 (quad permute [1, 0, 3, 2]; every lane of the wave must
// execute it).
__device__ inline double quad_xor1(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Global index of the tile's local amplitude j: the low c bits are contiguous,
// the rest sit at bit lo upward (lo < 0: a contiguous 4096 block). The map
// scatters bits independently, so the index of j = a | b (disjoint bits) is
// base + scat(a) + scat(b).
__device__ inline unsigned long long tile_base(long long tile, int c, int lo) {
    if (lo < 0) return (unsigned long long)tile << kTileBits;
    const int lo_span = lo - c;  // tile-id bits placed in [c, lo)
    const unsigned long long tl = (unsigned long long)tile & ((1ull << lo_span) - 1ull);
    const unsigned long long th = (unsigned long long)tile >> lo_span;
    return (tl << c) | (th << (lo + (kTileBits - c)));
}
__device__ inline unsigned long long scat(unsigned j, int c, int lo) {
    if (lo < 0) return j;
    return (unsigned long long)(j & ((1u << c) - 1u)) | ((unsigned long long)(j >> c) << lo);
}

// Per-tile energy tables (all 256 threads): tabE[0..127] = T1, T2; cst = C.
__device__ void build_energy(const Term* tm, const int* off, unsigned long long base, double* tabE, double* cst) {
    const int t = threadIdx.x;
    if (t < 128) {
        const bool lowh = t < 64;
        const unsigned u = (unsigned)(t & 63);
        double v = 0.0;
        const int kp = lowh ? kLL : kHH, kb = lowh ? kBL : kBH, kf = lowh ? kFL : kFH;
        for (int q = off[kp]; q < off[kp + 1]; ++q) v -= tm[q].w * bit_spin(u, tm[q].a) * bit_spin(u, tm[q].b);
        for (int q = off[kb]; q < off[kb + 1]; ++q) v -= tm[q].w * bit_spin(base, tm[q].a) * bit_spin(u, tm[q].b);
        for (int q = off[kf]; q < off[kf + 1]; ++q) v -= tm[q].w * bit_spin(u, tm[q].a);
        tabE[t] = v;
    } else if (t < 192) {
        const int lane = t & 63;
        double v = 0.0;
        for (int q = off[kBB] + lane; q < off[kBB + 1]; q += 64)
            v -= tm[q].w * bit_spin(base, tm[q].a) * bit_spin(base, tm[q].b);
        for (int q = off[kFB] + lane; q < off[kFB + 1]; q += 64) v -= tm[q].w * bit_spin(base, tm[q].a);
        v = wave_sum(v);
        if (lane == 0) *cst = v;
    }
}

// Phasor tables exp(-i scale E) for the energy tables, the tile constant and
// exp(+i scale w) of every cross edge (after build_energy + barrier).
__device__ void build_phasors(const Term* tm, const int* off, double scale, const double* tabE, double cst,
                              double2* tabP, double2* crossP, double2* cstP) {
    const int t = threadIdx.x;
    if (t < 128) {
        double sn, cs;
        sincos(scale * tabE[t], &sn, &cs);
        tabP[t] = make_double2(cs, -sn);
    } else if (t == 128) {
        double sn, cs;
        sincos(scale * cst, &sn, &cs);
        *cstP = make_double2(cs, -sn);
    } else if (t >= 192) {
        for (int q = t - 192; q < off[kLH + 1] - off[kLH]; q += 64) {
            double sn, cs;
            sincos(scale * tm[off[kLH] + q].w, &sn, &cs);
            crossP[q] = make_double2(cs, sn);  // factor when s_a s_b = +1
        }
    }
}

__device__ inline double tile_energy(const Term* tm, const int* off, int j, const double* tabE, double cst) {
    double e = cst + tabE[j & 63] + tabE[64 + (j >> 6)];
    for (int q = off[kLH]; q < off[kLH + 1]; ++q) {
        const bool same = (((j >> tm[q].a) ^ (j >> tm[q].b)) & 1) == 0;
        e += same ? -tm[q].w : tm[q].w;
    }
    return e;
}

__device__ inline double2 tile_phasor(const Term* tm, const int* off, int j, const double2* tabP, const double2* crossP,
                                      double2 cstP) {
    double2 f = cmul(cstP, cmul(tabP[j & 63], tabP[64 + (j >> 6)]));
    for (int q = off[kLH]; q < off[kLH + 1]; ++q) {
        const bool same = (((j >> tm[q].a) ^ (j >> tm[q].b)) & 1) == 0;
        const double2 x = crossP[q - off[kLH]];
        f = cmul(f, same ? x : cconj(x));
    }
    return f;
}

// Per-tile phase data of a pass with the register phase block, written by
// k_sv_tile_phases just before the pass so that the pass evaluates no
// transcendental: the base-local field f_a on each local bit a (the sum of
// w * s_b over the couplings of bit a to base bits b; bits 0-5 are L, 6-11 H),
// the tile constant C (base-base couplings and base fields), exp(+i s_c f_a),
// exp(+i s_out f_a) of the L bits and exp(-i s_c C), where s_c = s_out + s_in
// is the scale of the pass's one combined phase.
struct alignas(16) TileRec {
    double f[12];
    double cst;
    double pad;
    c2 g[12];
    c2 gl[6];
    c2 cp;
};
static_assert(sizeof(TileRec) == 416, "tile record layout");
constexpr int kRecChunks = (int)(sizeof(TileRec) / 16);  // 26 16-byte pieces

__global__ void __launch_bounds__(256) k_sv_tile_phases(PassTerms pt, int c, int lo, long long ntiles, double s_c,
                                                        double s_out, TileRec* __restrict__ rec) {
    const long long tile = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= ntiles) return;
    const unsigned long long base = tile_base(tile, c, lo);
    const Term* tm = pt.t;
    TileRec r;
#pragma unroll 1
    for (int a = 0; a < 12; ++a) {
        const int kb = a < 6 ? kBL : kBH, bit = a % 6;
        double f = 0.0;
        for (int q = pt.off[kb]; q < pt.off[kb + 1]; ++q)
            if (tm[q].b == bit) f += tm[q].w * bit_spin(base, tm[q].a);
        r.f[a] = f;
        double sn, cs;
        sincos(s_c * f, &sn, &cs);
        r.g[a] = c2{cs, sn};
        if (a < 6) {
            sincos(s_out * f, &sn, &cs);
            r.gl[a] = c2{cs, sn};
        }
    }
    double v = 0.0;
    for (int q = pt.off[kBB]; q < pt.off[kBB + 1]; ++q) v -= tm[q].w * bit_spin(base, tm[q].a) * bit_spin(base, tm[q].b);
    for (int q = pt.off[kFB]; q < pt.off[kFB + 1]; ++q) v -= tm[q].w * bit_spin(base, tm[q].a);
    r.cst = v;
    r.pad = 0.0;
    double sn, cs;
    sincos(s_c * v, &sn, &cs);
    r.cp = c2{cs, -sn};
    rec[tile] = r;
}

// quantum.rs:196-203 on register amplitudes: (a, b) = (x, x|bit)
__device__ inline void rotate_pair(c2& a, c2& b, double c, double s) {
    const double ra = a.x, ia = a.y, rb = b.x, ib = b.y;
    a = c2{c * ra - s * ib, c * ia + s * rb};
    b = c2{c * rb - s * ia, c * ib + s * ra};
}

// Rotate the register bits selected by `mask4` (bit i of mask4 <-> register bit i).
__device__ inline void rotate_regs(c2 (&v)[kRegs], unsigned mask4, double c, double s) {
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        if (!((mask4 >> bit) & 1u)) continue;
#pragma unroll
        for (int r = 0; r < kRegs; ++r)
            if (!(r & (1 << bit))) rotate_pair(v[r], v[r | (1 << bit)], c, s);
    }
}

// Register layouts of a tile (lay0, lay1, lay2 above).
enum { kLay0 = 0, kLay1 = 1, kLay2 = 2 };
__device__ inline int lay_of(int which, int t, int r) {
    return which == kLay0 ? lay0(t, r) : which == kLay1 ? lay1(t, r) : lay2(t, r);
}

// One full-tile exchange through LDS from layout FROM to layout TO; `mid` runs
// between the two barriers (it sees every LDS write made before the exchange).
template <int FROM, int TO, class Mid>
__device__ inline void exchange(c2 (&v)[kRegs], c2* sh, Mid&& mid) {
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < kRegs; ++r) sh[sidx(lay_of(FROM, t, r))] = v[r];
    __syncthreads();
    mid();
#pragma unroll
    for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay_of(TO, t, r))];
    __syncthreads();
}

// Rotates the group bits of a tile with two exchanges: from lay0 (registers hold
// bits 8-11) through lay2 (bits 0-3) to lay1 (bits 4-7), or back from lay1 to
// lay0. The qubits of a group commute; within a group the order differs from
// the upstream ascending loop only by rounding.
template <class Mid>
__device__ inline void rotate_group_fwd(c2 (&v)[kRegs], c2* sh, unsigned mask, double c, double s, Mid&& mid) {
    rotate_regs(v, (mask >> 8) & 15u, c, s);
    exchange<kLay0, kLay2>(v, sh, mid);
    rotate_regs(v, mask & 15u, c, s);
    exchange<kLay2, kLay1>(v, sh, [] {});
    rotate_regs(v, (mask >> 4) & 15u, c, s);
}
__device__ inline void rotate_group_bwd(c2 (&v)[kRegs], c2* sh, unsigned mask, double c, double s) {
    rotate_regs(v, (mask >> 4) & 15u, c, s);
    exchange<kLay1, kLay2>(v, sh, [] {});
    rotate_regs(v, mask & 15u, c, s);
    exchange<kLay2, kLay0>(v, sh, [] {});
    rotate_regs(v, (mask >> 8) & 15u, c, s);
}

// One streaming pass. PH: phase block (merged pass); RA / RB: rotate before /
// after it; FAST: the phase block works on the registers from per-tile tables
// (k_sv_tile_phases) and pass-constant coupling tables, else on the LDS copy of
// the tile from tables built in the pass. The tile is loaded in lay0; a rotation
// from lay0 ends in lay1 and one from lay1 in lay0, so the phase layout and the
// store layout follow from RA and RB at compile time.
template <bool PH, bool RA, bool RB, bool FAST>
__global__ void __launch_bounds__(kThreads, 2) k_sv_pass(double2* __restrict__ psi, PassTerms pt, PassArgs pa,
                                                      long long ntiles, const TileRec* __restrict__ rec,
                                                      double* __restrict__ partials, double* __restrict__ best_p,
                                                      long long* __restrict__ best_x) {
    constexpr int cur = RA ? kLay1 : kLay0;                             // layout of the phase block
    constexpr int out_lay = RB ? (cur == kLay1 ? kLay0 : kLay1) : cur;  // layout of the store
    __shared__ c2 sh[kTile + kTile / 16];
    __shared__ Term tm[FAST ? 1 : kMaxTerms];  // the LDS phase block's terms (FAST reads pt.t once)
    __shared__ int off[kKinds + 1];
    __shared__ double tabE[128];
    __shared__ double2 tabPo[128], tabPi[FAST ? 64 : 128];
    __shared__ double2 crossPo[kMaxCrossT], crossPi[FAST ? 1 : kMaxCrossT];
    __shared__ double cst;
    __shared__ double2 cstPo, cstPi;
    __shared__ double red[kThreads / 64][kObs + 1];
    __shared__ long long redx[kThreads / 64];
    __shared__ double xE[1 << kFastCross];
    __shared__ c2 xCorr[1 << kFastCross];  // exp(-i s_out e(k)) conj(exp(-i s_out e(k ^ x0mask)))
    __shared__ c2 shrec[2][kRecChunks];  // by tile parity: a pass without RB reads it after its last barrier
    const int t = threadIdx.x;
    // one combined phase covers phase_out and phase_in
    const double s_c = (pa.phase_out ? pa.scale_out : 0.0) + (pa.phase_in ? pa.scale_in : 0.0);
    if (PH) {
        if (!FAST)
            for (int q = t; q < pt.off[kKinds]; q += kThreads) tm[q] = pt.t[q];
        if (t <= kKinds) off[t] = pt.off[t];
        if (FAST && t < (1 << pa.nx)) {
            // pass constants: energy and phasors of the L-H couplings per sign pattern
            // (bit q of t set <-> the spins of coupling q differ: -w s s = +w), and the
            // rho01 factor exp(-i s_out e(k)) conj(exp(-i s_out e(k ^ x0mask)))
            auto ecross = [&](int k) {
                double e = 0.0;
                for (int q = 0; q < pa.nx; ++q) e += ((k >> q) & 1) ? pa.xw[q] : -pa.xw[q];
                return e;
            };
            const double e = ecross(t), e1 = ecross(t ^ pa.x0mask);
            xE[t] = e;
            double sn, cs, sn1, cs1;
            sincos(s_c * e, &sn, &cs);
            crossPo[t] = make_double2(cs, -sn);  // exp(-i s_c e)
            sincos(pa.scale_out * e, &sn, &cs);
            sincos(pa.scale_out * e1, &sn1, &cs1);
            xCorr[t] = cm(c2{cs, -sn}, c2{cs1, sn1});
        }
        __syncthreads();
    }
    // FAST: pass constants kept in registers across tiles: thread t < 128 holds the
    // base-independent part Tc of table entry t (LL edges and L fields for t < 64,
    // HH edges and H fields for t >= 64) and its phasor exp(-i s_c Tc); threads
    // 128..191 hold exp(-i s_out Tc) of the L entries.
    double tc = 0.0;
    c2 pcc{1.0, 0.0};
    if (PH && FAST && t < 192) {
        const int u = t < 128 ? (t & 63) : t - 128;
        const bool lowh = t < 64 || t >= 128;
        const int kp = lowh ? kLL : kHH, kf = lowh ? kFL : kFH;
        const Term* T = pt.t;
        for (int q = pt.off[kp]; q < pt.off[kp + 1]; ++q) tc -= T[q].w * bit_spin(u, T[q].a) * bit_spin(u, T[q].b);
        for (int q = pt.off[kf]; q < pt.off[kf + 1]; ++q) tc -= T[q].w * bit_spin(u, T[q].a);
        double sn, cs;
        sincos((t < 128 ? s_c : pa.scale_out) * tc, &sn, &cs);
        pcc = c2{cs, -sn};
    }
    const int c = pa.c, lo = pa.lo;
    // per-thread parts of the load and store addresses (the register parts are uniform)
    const unsigned long long toff_ld = scat((unsigned)lay_of(kLay0, t, 0), c, lo);
    const unsigned long long toff_st = scat((unsigned)lay_of(out_lay, t, 0), c, lo);
    double acc[kObs] = {0, 0, 0, 0, 0, 0};
    Best best{-1.0, 0};

    auto load_tile = [&](c2 (&w)[kRegs], c2& wr, long long tile) __attribute__((always_inline)) {
        const double2* src = psi + tile_base(tile, c, lo) + toff_ld;
#pragma unroll
        for (int r = 0; r < kRegs; ++r) w[r] = ldc(src + scat((unsigned)lay_of(kLay0, 0, r), c, lo));
        // every thread loads a piece of the record (threads >= kRecChunks repeat the
        // last one) so that no load sits under a divergent branch
        if (PH && FAST) wr = ldc(reinterpret_cast<const double2*>(rec + tile) + min(t, kRecChunks - 1));
    };

    // FAST table build from the tile record (threads < 192, after the record is in LDS)
    auto build_tables = [&](int par) __attribute__((always_inline)) {
        const TileRec* R = reinterpret_cast<const TileRec*>(shrec[par]);
        if (t < 128) {
            const int u = t & 63, half = t < 64 ? 0 : 6;
            double e = tc;
            c2 ph = pcc;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const bool up = (u >> a) & 1;
                e -= up ? R->f[half + a] : -R->f[half + a];
                ph = cm(ph, up ? R->g[half + a] : cj(R->g[half + a]));
            }
            tabE[t] = e;
            tabPo[t] = make_double2(ph.x, ph.y);  // exp(-i s_c T)
        } else if (t < 192) {
            const int u = t - 128;
            c2 ph = pcc;
#pragma unroll
            for (int a = 0; a < 6; ++a) ph = cm(ph, ((u >> a) & 1) ? R->gl[a] : cj(R->gl[a]));
            tabPi[u] = make_double2(ph.x, ph.y);  // exp(-i s_out T1)
        }
    };

    // One tile, loaded into v (and its record piece into vr); `par` alternates
    // between consecutive tiles of the workgroup.
    auto body = [&](c2 (&v)[kRegs], c2& vr, long long tile, int par) __attribute__((always_inline)) {
        const unsigned long long base = tile_base(tile, c, lo);
        if (PH && FAST) {
            if (t < kRecChunks) shrec[par][t] = vr;
            if (!RA) {
                __syncthreads();
                build_tables(par);
                __syncthreads();
            }
        }
        if (PH && !FAST) {
            build_energy(tm, off, base, tabE, &cst);
            __syncthreads();
            if (pa.phase_out) build_phasors(tm, off, pa.scale_out, tabE, cst, tabPo, crossPo, &cstPo);
            if (pa.phase_in) build_phasors(tm, off, pa.scale_in, tabE, cst, tabPi, crossPi, &cstPi);
            __syncthreads();
        }
        if (RA) {
            if (PH && FAST)
                rotate_group_fwd(v, sh, pa.mask, pa.ca, pa.sa, [&] { build_tables(par); });  // lay0 -> lay1
            else
                rotate_group_fwd(v, sh, pa.mask, pa.ca, pa.sa, [] {});
        }
        if (PH && FAST && cur != kLay1) {
            // phase block, one sweep: E(j) = C + T1[j & 63] + T2[j >> 6] + xE[k(j)]
            // with k(j) the sign pattern of the L-H couplings; one phasor
            // exp(-i s_c E) applies phase_out and phase_in together. rho01 pairs
            // (j, j ^ 1) sit in lanes (t, t ^ 1) of the same register slot (bit 0
            // of j is bit 0 of t in lay0 and lay1) and need the phase_out factor
            // Po(j) conj(Po(j ^ 1)), in which C and T2 cancel:
            // P1o[lo] conj(P1o[lo ^ 1]) xPo[k] conj(xPo[k ^ x0mask]). The tile goes
            // through each thread's own LDS slots (no barrier: a thread reads back
            // only what it wrote, and the next exchange writes the same slots) so
            // the loop is unrolled only four ways. (This form serves the first pass
            // of an anneal, in lay0; the passes in lay1 take the block below.)
            const TileRec* R = reinterpret_cast<const TileRec*>(shrec[par]);
            const double cstv = R->cst;
            const c2 cp = R->cp;
            const bool even = (t & 1) == 0;
#pragma unroll
            for (int r = 0; r < kRegs; ++r) sh[sidx(lay_of(cur, t, r))] = v[r];
#pragma unroll 4
            for (int r = 0; r < kRegs; ++r) {
                const int j = lay_of(cur, t, r);
                int k = 0;
#pragma unroll
                for (int q = 0; q < kFastCross; ++q)
                    if (q < pa.nx) k |= (((j >> pa.xa[q]) ^ (j >> pa.xb[q])) & 1) << q;
                const int jl = j & 63, jh = j >> 6;
                const c2 u = sh[sidx(j)];
                if (pa.phase_out) {
                    const double p = u.x * u.x + u.y * u.y;
                    const double e = cstv + tabE[jl] + tabE[64 + jh] + xE[k];
                    acc[4] += p;
                    acc[5] += p * e;
                    if (pa.best) best = better(best, Best{p, (long long)(base + scat((unsigned)j, c, lo))});
                    const c2 wv{__shfl_xor(u.x, 1, 64), __shfl_xor(u.y, 1, 64)};
                    if (even) {
                        const c2 r1 = cm(c2of(tabPi[jl]), cj(c2of(tabPi[jl ^ 1])));
                        const c2 corr = cm(r1, xCorr[k]);
                        const c2 uw{u.x * wv.x + u.y * wv.y, u.y * wv.x - u.x * wv.y};
                        const c2 z = cm(uw, corr);
                        acc[0] += p;
                        acc[1] += wv.x * wv.x + wv.y * wv.y;
                        acc[2] += z.x;
                        acc[3] += z.y;
                    }
                }
                const c2 ph = cm(cm(cp, c2of(tabPo[jl])), cm(c2of(tabPo[64 + jh]), c2of(crossPo[k])));
                sh[sidx(j)] = cm(u, ph);
            }
#pragma unroll
            for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay_of(cur, t, r))];
        }
        if (PH && FAST && cur == kLay1) {
            // The same phase block in lay1, j = (t & 15) | (r << 4) | ((t >> 4) << 8):
            // with r = 4b + a the L index jl = (t & 15) | (a << 4) takes four values
            // per thread and the H index jh = b | ((t >> 4) << 2) four more, so their
            // table entries (energy, combined phasor, rho01 factor of the L part) are
            // read once per tile; the coupling part comes from one pass-constant
            // entry per sign pattern. The rho01 partner j ^ 1 is lane t ^ 1, read
            // with a quad permute.
            const TileRec* R = reinterpret_cast<const TileRec*>(shrec[par]);
            const double cstv = R->cst;
            const c2 cp = R->cp;
            const bool even = (t & 1) == 0;
            // k(j) = k(T) ^ k(R) for j = T | R with disjoint bits: the thread part once,
            // the register part (uniform) per slot
            auto kpat = [&](int j) {
                int k = 0;
#pragma unroll
                for (int q = 0; q < kFastCross; ++q)
                    if (q < pa.nx) k |= (((j >> pa.xa[q]) ^ (j >> pa.xb[q])) & 1) << q;
                return k;
            };
            const int kt = kpat(lay_of(kLay1, t, 0));
            double EL[4];
            c2 PL[4], R1[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int jl = (t & 15) | (a << 4);
                EL[a] = cstv + tabE[jl];
                PL[a] = cm(cp, c2of(tabPo[jl]));
                R1[a] = cm(c2of(tabPi[jl]), cj(c2of(tabPi[jl ^ 1])));
            }
#pragma unroll
            for (int r = 0; r < kRegs; ++r) sh[sidx(lay_of(kLay1, t, r))] = v[r];
#pragma unroll 1
            for (int b = 0; b < 4; ++b) {
                const int jh = b | ((t >> 4) << 2);
                const double eh = tabE[64 + jh];
                const c2 phb = c2of(tabPo[64 + jh]);
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int j = lay_of(kLay1, t, 4 * b + a);
                    const int k = kt ^ kpat(lay_of(kLay1, 0, 4 * b + a));
                    const c2 u = sh[sidx(j)];
                    if (pa.phase_out) {
                        const double p = u.x * u.x + u.y * u.y;
                        const double e = EL[a] + eh + xE[k];
                        acc[4] += p;
                        acc[5] += p * e;
                        if (pa.best) best = better(best, Best{p, (long long)(base + scat((unsigned)j, c, lo))});
                        const c2 wv{quad_xor1(u.x), quad_xor1(u.y)};
                        if (even) {
                            const c2 corr = cm(R1[a], xCorr[k]);
                            const c2 uw{u.x * wv.x + u.y * wv.y, u.y * wv.x - u.x * wv.y};
                            const c2 z = cm(uw, corr);
                            acc[0] += p;
                            acc[1] += wv.x * wv.x + wv.y * wv.y;
                            acc[2] += z.x;
                            acc[3] += z.y;
                        }
                    }
                    sh[sidx(j)] = cm(u, cm(cm(PL[a], phb), c2of(crossPo[k])));
                }
            }
#pragma unroll
            for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay_of(kLay1, t, r))];
        }
        if (PH && !FAST) {
            // phase block on the LDS copy of the tile (lay0 order), plain loops keep registers low
#pragma unroll
            for (int r = 0; r < kRegs; ++r) sh[sidx(lay_of(cur, t, r))] = v[r];
            __syncthreads();
            if (pa.phase_out) {
#pragma unroll 1
                for (int j = t; j < kTile; j += kThreads) {
                    const double2 f = tile_phasor(tm, off, j, tabPo, crossPo, cstPo);
                    sh[sidx(j)] = cm(sh[sidx(j)], c2of(f));
                }
                __syncthreads();
#pragma unroll 1
                for (int j = t; j < kTile; j += kThreads) {
                    const c2 u = sh[sidx(j)];
                    const double e = tile_energy(tm, off, j, tabE, cst);
                    const double p = u.x * u.x + u.y * u.y;
                    acc[4] += p;
                    acc[5] += p * e;
                    if (pa.best) best = better(best, Best{p, (long long)(base + scat((unsigned)j, c, lo))});
                    if ((j & 1) == 0) {  // local bit 0 == global bit 0
                        const c2 w2 = sh[sidx(j + 1)];
                        acc[0] += u.x * u.x + u.y * u.y;
                        acc[1] += w2.x * w2.x + w2.y * w2.y;
                        acc[2] += u.x * w2.x + u.y * w2.y;
                        acc[3] += u.y * w2.x - u.x * w2.y;
                    }
                }
                __syncthreads();
            }
            if (pa.phase_in) {
#pragma unroll 1
                for (int j = t; j < kTile; j += kThreads) {
                    const double2 f = tile_phasor(tm, off, j, tabPi, crossPi, cstPi);
                    sh[sidx(j)] = cm(sh[sidx(j)], c2of(f));
                }
                __syncthreads();
            }
#pragma unroll
            for (int r = 0; r < kRegs; ++r) v[r] = sh[sidx(lay_of(cur, t, r))];
            __syncthreads();
        }
        if (RB) {
            if (cur == kLay1)
                rotate_group_bwd(v, sh, pa.mask, pa.cb, pa.sb);
            else
                rotate_group_fwd(v, sh, pa.mask, pa.cb, pa.sb, [] {});
        }
        double2* dst = psi + base + toff_st;
#pragma unroll
        for (int r = 0; r < kRegs; ++r) stc(dst + scat((unsigned)lay_of(out_lay, 0, r), c, lo), v[r]);
        if (PH && !FAST) __syncthreads();  // tables are rebuilt for the next tile
    };

    int par = 0;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x, par ^= 1) {
        c2 v[kRegs];
        c2 vr{0.0, 0.0};
        load_tile(v, vr, tile);
        body(v, vr, tile, par);
    }
    if (PH && pa.phase_out) {
#pragma unroll
        for (int q = 0; q < kObs; ++q) acc[q] = wave_sum(acc[q]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) best = better(best, Best{__shfl_xor(best.p, o, 64), __shfl_xor(best.x, o, 64)});
        if ((t & 63) == 0) {
            for (int q = 0; q < kObs; ++q) red[t >> 6][q] = acc[q];
            red[t >> 6][kObs] = best.p;
            redx[t >> 6] = best.x;
        }
        __syncthreads();
        if (t == 0) {
            Best b{-1.0, 0};
            for (int q = 0; q < kObs; ++q) {
                double s = 0.0;
                for (int w = 0; w < kThreads / 64; ++w) s += red[w][q];
                partials[(long long)blockIdx.x * kObs + q] = s;
            }
            for (int w = 0; w < kThreads / 64; ++w) b = better(b, Best{red[w][kObs], redx[w]});
            if (pa.best) {
                best_p[blockIdx.x] = b.p;
                best_x[blockIdx.x] = b.x;
            }
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

// Final measurement statistics of the resident path (quantum.rs:89-102): total,
// sum p*E, and the argmax with the lowest index among equal maxima.

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

SvOutcome sv_anneal_device(const SvProblem& pb, bool keep_state, DevArena* arena) {
    DevArena own;
    DevArena& ar = arena ? *arena : own;
    const int n = pb.n;
    if (n < 1 || n > 40) throw Error(-1, "state-vector anneal needs 1 <= n <= 40");
    if (pb.steps < 1) throw Error(-1, "state-vector anneal needs steps >= 1");
    const long long dim = 1ll << n;
    const int m = (int)pb.ew.size();
    SvOutcome out;
    out.entropy.assign(pb.steps, 0.0);
    out.energy_trace.assign(pb.steps, 0.0);

    // model upload
    int* d_ei = ar.take<int>(std::max(m, 1));
    int* d_ej = ar.take<int>(std::max(m, 1));
    double* d_ew = ar.take<double>(std::max(m, 1));
    double* d_h = ar.take<double>(n);
    if (m) {
        TQ_HIP(hipMemcpy(d_ei, pb.ei.data(), m * sizeof(int), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_ej, pb.ej.data(), m * sizeof(int), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_ew, pb.ew.data(), m * sizeof(double), hipMemcpyHostToDevice));
    }
    TQ_HIP(hipMemcpy(d_h, pb.h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    Model md{n, m, d_ei, d_ej, d_ew, d_h};

    // per-step rotation cos/sin from the host libm, as quantum.rs:191
    std::vector<double> rc(pb.steps), rsn(pb.steps);
    for (int k = 0; k < pb.steps; ++k) {
        const double th = pb.a[k] * pb.dt;
        rc[k] = std::cos(th);
        rsn[k] = std::sin(th);
    }
    const double amp0 = 1.0 / std::sqrt((double)dim);  // quantum.rs:72

    double2* psi = ar.take<double2>((std::size_t)dim);
    double* d_ent = ar.take<double>(pb.steps);
    double* d_en = ar.take<double>(pb.steps);
    hipStream_t st = nullptr;

    double total = 0.0, expected = 0.0, best_p = -1.0;
    long long best = 0;
    if (n <= kResidentMaxN) {
        double* d_b = ar.take<double>(pb.steps);
        double* d_c = ar.take<double>(pb.steps);
        double* d_s = ar.take<double>(pb.steps);
        TQ_HIP(hipMemcpy(d_b, pb.b.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_c, rc.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(d_s, rsn.data(), pb.steps * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_sv_resident, dim3(1), dim3(1024), 0, st, md, d_b, d_c, d_s, pb.dt,
                           pb.steps, amp0, d_ent, d_en, psi);
        TQ_LAUNCH_CHECK();
        const int mblocks = (int)std::min<long long>((dim + 255) / 256, 2048);
        double* ps = ar.take<double>((std::size_t)mblocks * 3);
        long long* pi = ar.take<long long>(mblocks);
        hipLaunchKernelGGL(k_sv_measure, dim3(mblocks), dim3(256), 0, st, psi, md, dim, ps, pi);
        TQ_LAUNCH_CHECK();
        TQ_HIP(hipStreamSynchronize(st));
        std::vector<double> hs(mblocks * 3);
        std::vector<long long> hi(mblocks);
        TQ_HIP(hipMemcpy(hs.data(), ps, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
        TQ_HIP(hipMemcpy(hi.data(), pi, hi.size() * sizeof(long long), hipMemcpyDeviceToHost));
        for (int b = 0; b < mblocks; ++b) {
            total += hs[b * 3 + 0];
            expected += hs[b * 3 + 1];
            if (hs[b * 3 + 2] > best_p || (hs[b * 3 + 2] == best_p && hi[b] < best)) {
                best_p = hs[b * 3 + 2];
                best = hi[b];
            }
        }
    } else {
        const long long ntiles = dim >> kTileBits;
        // 4096 workgroups of 64 tiles each at n = 30 (measured: one resident grid of two
        // workgroups per CU streams 10% slower, 16384 workgroups the same)
        const int nblocks = (int)std::min<long long>(ntiles, 4096);
        const bool fast_phase = true;  // register phase block wherever the L-H couplings fit
        double* partials = ar.take<double>((std::size_t)nblocks * kObs);
        double* bp = ar.take<double>(nblocks);
        long long* bx = ar.take<long long>(nblocks);
        hipLaunchKernelGGL(k_sv_fill, dim3(2048), dim3(256), 0, st, psi, dim, amp0);
        TQ_LAUNCH_CHECK();
        // rotation groups: bits [0,12) as contiguous blocks, then windows of <= 9 bits
        struct Group {
            int c, lo;
            unsigned mask;
        };
        std::vector<Group> groups;
        groups.push_back({kTileBits, -1, (1u << kTileBits) - 1u});
        for (int lo = kTileBits; lo < n; lo += 9) {
            const int w = std::min(9, n - lo);
            const int c = kTileBits - w;
            groups.push_back({c, lo, ((1u << w) - 1u) << c});
        }
        const int ng = (int)groups.size();
        // energy terms of every group, classified by tile position
        std::vector<Term> terms;
        std::vector<PassTerms> pts(ng);
        std::vector<std::size_t> starts(ng);
        for (int g = 0; g < ng; ++g) {
            std::vector<int> role(n, -1);
            for (int q = 0; q < kTileBits; ++q) {
                const int gb = g == 0 ? q : (q < groups[g].c ? q : groups[g].lo + q - groups[g].c);
                if (gb < n) role[gb] = q;
            }
            std::vector<Term> bucket[kKinds];
            auto cls = [](int q) { return q < 0 ? 0 : (q < 6 ? 1 : 2); };  // B, L, H
            for (int k = 0; k < m; ++k) {
                const int gi = pb.ei[k], gj = pb.ej[k];
                const int qi = role[gi], qj = role[gj];
                const int ci = cls(qi), cj = cls(qj);
                const double w = pb.ew[k];
                if (ci == 1 && cj == 1) bucket[kLL].push_back({qi, qj, w});
                else if (ci == 2 && cj == 2) bucket[kHH].push_back({qi - 6, qj - 6, w});
                else if (ci == 0 && cj == 0) bucket[kBB].push_back({gi, gj, w});
                else if (ci == 0 && cj == 1) bucket[kBL].push_back({gi, qj, w});
                else if (ci == 1 && cj == 0) bucket[kBL].push_back({gj, qi, w});
                else if (ci == 0 && cj == 2) bucket[kBH].push_back({gi, qj - 6, w});
                else if (ci == 2 && cj == 0) bucket[kBH].push_back({gj, qi - 6, w});
                else bucket[kLH].push_back({qi, qj, w});
            }
            for (int i = 0; i < n; ++i) {
                const int q = role[i];
                if (cls(q) == 1) bucket[kFL].push_back({q, 0, pb.h[i]});
                else if (cls(q) == 2) bucket[kFH].push_back({q - 6, 0, pb.h[i]});
                else bucket[kFB].push_back({i, 0, pb.h[i]});
            }
            if (bucket[kLH].size() > (std::size_t)kMaxCrossT)
                throw Error(-1, "too many couplings between the two halves of a tile");
            // base-local couplings grouped by local bit (the per-bit fields of the tables)
            for (int kd : {kBL, kBH})
                std::stable_sort(bucket[kd].begin(), bucket[kd].end(),
                                 [](const Term& x, const Term& y) { return x.b < y.b; });
            starts[g] = terms.size();
            int acc = 0;
            for (int kd = 0; kd < kKinds; ++kd) {
                pts[g].off[kd] = acc;
                acc += (int)bucket[kd].size();
                terms.insert(terms.end(), bucket[kd].begin(), bucket[kd].end());
            }
            pts[g].off[kKinds] = acc;
            if (acc > kMaxTerms) throw Error(-1, "more than 256 energy terms (m + n) per tile");
        }
        TileRec* d_rec = ar.take<TileRec>((std::size_t)ntiles);
        Term* d_terms = ar.take<Term>(std::max<std::size_t>(terms.size(), 1));
        if (!terms.empty())
            TQ_HIP(hipMemcpy(d_terms, terms.data(), terms.size() * sizeof(Term), hipMemcpyHostToDevice));
        for (int g = 0; g < ng; ++g) pts[g].t = d_terms + starts[g];

        auto launch = [&](int g, const PassArgs& pa, bool ph) {
            // the passes the host loop below issues: rotate only; phase then rotate (first);
            // rotate, phase, rotate (step boundary); rotate then phase (last)
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kThreads), 0, st, psi, pts[g], pa, ntiles, d_rec,
                                   partials, bp, bx);
            };
            const bool f = pa.fast != 0;
            if (ph && f) {
                // the tile records of this pass (fields and phasors of every tile)
                const double s_c = (pa.phase_out ? pa.scale_out : 0.0) + (pa.phase_in ? pa.scale_in : 0.0);
                const int rb = (int)((ntiles + 255) / 256);
                hipLaunchKernelGGL(k_sv_tile_phases, dim3(rb), dim3(256), 0, st, pts[g], pa.c, pa.lo, ntiles, s_c,
                                   pa.scale_out, d_rec);
                TQ_LAUNCH_CHECK();
            }
            if (!ph && pa.rot_a && !pa.rot_b) go(k_sv_pass<false, true, false, false>);
            else if (ph && !pa.rot_a && pa.rot_b) f ? go(k_sv_pass<true, false, true, true>) : go(k_sv_pass<true, false, true, false>);
            else if (ph && pa.rot_a && pa.rot_b) f ? go(k_sv_pass<true, true, true, true>) : go(k_sv_pass<true, true, true, false>);
            else if (ph && pa.rot_a && !pa.rot_b) f ? go(k_sv_pass<true, true, false, true>) : go(k_sv_pass<true, true, false, false>);
            else throw Error(-1, "state-vector pass: unsupported rotation pattern");
            TQ_LAUNCH_CHECK();
        };
        auto base_args = [&](int g) {
            PassArgs pa{};
            pa.c = g == 0 ? 0 : groups[g].c;
            pa.lo = groups[g].lo;
            pa.mask = groups[g].mask;
            const int nlh = pts[g].off[kLH + 1] - pts[g].off[kLH];
            if (fast_phase && nlh <= kFastCross) {
                pa.fast = 1;
                pa.nx = nlh;
                for (int q = 0; q < nlh; ++q) {
                    const Term& tq = terms[starts[g] + pts[g].off[kLH] + q];
                    pa.xa[q] = tq.a;
                    pa.xb[q] = tq.b;
                    pa.xw[q] = tq.w;
                    if (tq.a == 0 || tq.b == 0) pa.x0mask |= 1 << q;
                }
            }
            return pa;
        };
        for (int k = 0; k < pb.steps; ++k) {
            // ascending group order on even steps, descending on odd steps
            auto grp = [&](int step, int idx) { return (step % 2 == 0) ? idx : ng - 1 - idx; };
            for (int idx = 0; idx < ng; ++idx) {
                const int g = grp(k, idx);
                if (idx == 0 && k > 0) continue;  // rotated at the end of the previous step
                PassArgs pa = base_args(g);
                if (idx == 0) {  // very first pass: phase(0), rotate(0)
                    pa.phase_in = 1;
                    pa.scale_in = pb.b[0] * pb.dt * 0.5;
                    pa.rot_b = 1;
                    pa.cb = rc[0];
                    pa.sb = rsn[0];
                    launch(g, pa, true);
                } else if (idx < ng - 1) {
                    pa.rot_a = 1;
                    pa.ca = rc[k];
                    pa.sa = rsn[k];
                    launch(g, pa, false);
                } else {
                    pa.rot_a = 1;
                    pa.ca = rc[k];
                    pa.sa = rsn[k];
                    pa.phase_out = 1;
                    pa.scale_out = pb.b[k] * pb.dt * 0.5;
                    if (k + 1 < pb.steps) {
                        pa.phase_in = 1;
                        pa.scale_in = pb.b[k + 1] * pb.dt * 0.5;
                        pa.rot_b = 1;
                        pa.cb = rc[k + 1];
                        pa.sb = rsn[k + 1];
                    } else {
                        pa.best = 1;
                    }
                    launch(g, pa, true);
                    hipLaunchKernelGGL(k_sv_finish_step, dim3(1), dim3(256), 0, st, partials, nblocks, n, k,
                                       d_ent, d_en);
                    TQ_LAUNCH_CHECK();
                }
            }
        }
        TQ_HIP(hipStreamSynchronize(st));
        std::vector<double> hs((std::size_t)nblocks * kObs), hbp(nblocks);
        std::vector<long long> hbx(nblocks);
        TQ_HIP(hipMemcpy(hs.data(), partials, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
        TQ_HIP(hipMemcpy(hbp.data(), bp, nblocks * sizeof(double), hipMemcpyDeviceToHost));
        TQ_HIP(hipMemcpy(hbx.data(), bx, nblocks * sizeof(long long), hipMemcpyDeviceToHost));
        for (int b = 0; b < nblocks; ++b) {
            total += hs[(std::size_t)b * kObs + 4];
            expected += hs[(std::size_t)b * kObs + 5];
            if (hbp[b] > best_p || (hbp[b] == best_p && hbx[b] < best)) {
                best_p = hbp[b];
                best = hbx[b];
            }
        }
    }
    TQ_HIP(hipStreamSynchronize(st));
    TQ_HIP(hipMemcpy(out.entropy.data(), d_ent, pb.steps * sizeof(double), hipMemcpyDeviceToHost));
    TQ_HIP(hipMemcpy(out.energy_trace.data(), d_en, pb.steps * sizeof(double), hipMemcpyDeviceToHost));
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
        TQ_HIP(hipMemcpy(out.state.data(), psi, dim * sizeof(double2), hipMemcpyDeviceToHost));
    }
    return out;
}


// ---- static scratch: tessera_sv_anneal without device allocation ----------
// One array per GPU, held until the process ends; every call on a GPU reuses
// it (nothing in it needs to start at zero: each call writes every slot it
// reads). Host code never passes the array's name to a kernel or a copy: its
// address on each device comes from hipGetSymbolAddress, taken once per device
// with that device current. Calls on one GPU are serialised by a per-device
// mutex inside this process.
static __device__ __align__(256) unsigned char scratch[kSvStaticBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu[kMaxDevices];
void* g_static_ptr[kMaxDevices] = {};
std::once_flag g_static_once[kMaxDevices];
}  // namespace

SvOutcome sv_anneal_static(const SvProblem& pb) {
    if (pb.n < 1 || pb.n > kSvStaticMaxN)
        throw Error(-2, "static-scratch anneal needs 1 <= n <= " + std::to_string(kSvStaticMaxN));
    if (pb.steps < 1 || pb.steps > kSvStaticMaxSteps)
        throw Error(-2, "static-scratch anneal needs 1 <= steps <= " + std::to_string(kSvStaticMaxSteps));
    if (pb.ew.size() > (std::size_t)kSvStaticMaxEdges)
        throw Error(-2, "static-scratch anneal takes at most " + std::to_string(kSvStaticMaxEdges) + " edges");
    int dev = 0;
    TQ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(-3, "device index out of range");
    hipError_t status = hipSuccess;
    std::call_once(g_static_once[dev], [&] {
        void* ptr = nullptr;
        status = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (status == hipSuccess) g_static_ptr[dev] = ptr;
    });
    if (status != hipSuccess) throw Error(-3, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(status));
    if (!g_static_ptr[dev]) throw Error(-3, "static scratch address unavailable on this device");
    std::lock_guard<std::mutex> lock(g_static_mu[dev]);
    DevArena ar(g_static_ptr[dev], kSvStaticBytes);
    return sv_anneal_device(pb, false, &ar);
}

}  // namespace tq
