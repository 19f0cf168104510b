// Message-passing guidance network on gfx950 (layer algebra in gnn.hpp).
//
// Mapping: the hidden width is 64, one wavefront per node (lane = feature) for
// the gather/scatter and elementwise work, and the dense node products
// (H W_sd, [H M] W_u and their transposes in the backward pass) on MFMA
// v_mfma_f32_16x16x16_bf16 with BF16 operands and FP32 accumulation. Weight
// gradients are reduced over the node axis by split-K partial products and a
// deterministic column-sum pass (no atomics: the same batch gives the same bits).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "gnn.hpp"

namespace tq {

namespace {

using GL = GnnLayout;
constexpr int F = GL::F;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// C[z] = A B over the K range [z*kchunk, min(K, (z+1)*kchunk)). 256 threads:
// four wavefronts, each one 16x16 output block of the 32x32 workgroup tile.
// MFMA 16x16x16 operand layout: lane l holds A(row l%16, k 4(l/16)..+3) and
// B(k 4(l/16)..+3, col l%16); the result D(4(l/16)+i, l%16) in acc[i].
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, long sam, long sak,
                                                   const uint16_t* __restrict__ B, long sbk, long sbn,
                                                   float* __restrict__ C, int M, int N, int K, int kchunk) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * 32 + (wave >> 1) * 16;
    const int n0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    const int row = m0 + (lane & 15), col = n0 + (lane & 15), kq = (lane >> 4) * 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + kq + i;
            const bool kin = k < kend;
            a[i] = (kin && row < M) ? (short)A[(long)row * sam + (long)k * sak] : (short)0;
            b[i] = (kin && col < N) ? (short)B[(long)k * sbk + (long)col * sbn] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float* Cz = C + (long)blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = m0 + (lane >> 4) * 4 + i;
        if (r < M && col < N) Cz[(long)r * N + col] = acc[i];
    }
}

// out[blockIdx.y * C + c] = sum of rows [blockIdx.y*rows, ...) of column c (FP64 accumulation).
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ X, int R, int C, int rows,
                                                float* __restrict__ out) {
    __shared__ double red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rg = threadIdx.x >> 6;
    const int r0 = blockIdx.y * rows, r1 = min(R, r0 + rows);
    double acc = 0.0;
    if (c < C)
        for (int r = r0 + rg; r < r1; r += 4) acc += X[(long)r * C + c];
    red[rg][threadIdx.x & 63] = acc;
    __syncthreads();
    if (rg == 0 && c < C)
        out[(long)blockIdx.y * C + c] = (float)(red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                                red[3][threadIdx.x]);
}

// H0 = relu(x W_in + b_in); writes the FP32 activation and the BF16 operand half of X0.
__global__ void k_embed_fwd(int nodes, const float* __restrict__ feat, const float* __restrict__ w,
                            float* __restrict__ pre, float* __restrict__ h, uint16_t* __restrict__ xb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes * F) return;
    const int v = idx / F, f = idx % F;
    const float p = feat[2 * v] * w[GL::W_IN + f] + feat[2 * v + 1] * w[GL::W_IN + F + f] + w[GL::B_IN + f];
    const float hv = fmaxf(p, 0.f);
    pre[idx] = p;
    h[idx] = hv;
    xb[(long)v * 2 * F + f] = to_bf16(hv);
}

// M_v = sum_u relu(P_u + Q_v + J_uv w_J + b_m) into the BF16 half X[:, 64:128].
__global__ __launch_bounds__(256) void k_msg_fwd(int nodes, int n0, const int32_t* __restrict__ adj_off,
                                                 const int32_t* __restrict__ adj_nbr, const float* __restrict__ adj_j,
                                                 const float* __restrict__ pq, const float* __restrict__ wl,
                                                 uint16_t* __restrict__ xb) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
    if (v >= nodes) return;
    const float q = pq[(long)v * 2 * F + F + f];
    const float wj = wl[GL::W_J + f], bm = wl[GL::B_M + f];
    float acc = 0.f;
    const int e1 = adj_off[n0 + v + 1];
    for (int e = adj_off[n0 + v]; e < e1; ++e) {
        const int u = adj_nbr[e] - n0;
        acc += fmaxf(pq[(long)u * 2 * F + f] + q + adj_j[e] * wj + bm, 0.f);
    }
    xb[(long)v * 2 * F + F + f] = to_bf16(acc);
}

// Z += b_u (in place, kept for the backward mask); H_l+1 = H_l + relu(Z).
__global__ void k_update_fwd(int nodes, float* __restrict__ z, const float* __restrict__ wl,
                             const float* __restrict__ h, float* __restrict__ h1, uint16_t* __restrict__ xb_next) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes * F) return;
    const int v = idx / F, f = idx % F;
    const float zz = z[idx] + wl[GL::B_U + f];
    z[idx] = zz;
    const float hv = h[idx] + fmaxf(zz, 0.f);
    h1[idx] = hv;
    if (xb_next) xb_next[(long)v * 2 * F + f] = to_bf16(hv);
}

// Per-node head z_i = tanh(H_i . w_z + c_z); with targets, the loss term, the
// seed gradient dH = dz' w_z and the parameter-gradient row [dz' H_i, dz'].
__global__ __launch_bounds__(256) void k_head_node(int nodes, int n0, const float* __restrict__ h,
                                                   const float* __restrict__ w, const float* __restrict__ tz,
                                                   float* __restrict__ zout, float* __restrict__ rows,
                                                   float* __restrict__ dh, float* __restrict__ loss, float inv_n) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
    if (v >= nodes) return;
    const float hv = h[(long)v * F + f];
    const double dot = wave_sum((double)hv * w[GL::W_Z + f]) + w[GL::C_Z];
    const float z = (float)std::tanh(dot);
    if (f == 0) zout[v] = z;
    if (!rows) return;
    const float t = tz[n0 + v];
    const float dp = 2.f * (z - t) * inv_n * (1.f - z * z);
    rows[(long)v * (F + 1) + f] = dp * hv;
    if (f == 0) {
        rows[(long)v * (F + 1) + F] = dp;
        loss[v] = (z - t) * (z - t) * inv_n;
    }
    dh[(long)v * F + f] = dp * w[GL::W_Z + f];
}

// Per-graph head s_g = sigmoid(mean_i H_i . w_s + c_s); with targets, the loss
// term, the gradient row [ds' pool, ds'] and dH_i += ds' w_s / n_g.
__global__ __launch_bounds__(256) void k_head_graph(int ng, int g0, const int32_t* __restrict__ node_off,
                                                    const float* __restrict__ h, const float* __restrict__ w,
                                                    const float* __restrict__ ts, float* __restrict__ sout,
                                                    float* __restrict__ rows, float* __restrict__ dh,
                                                    float* __restrict__ loss, float inv_b) {
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
    if (g >= ng) return;
    const int n0 = node_off[g0];
    const int a = node_off[g0 + g] - n0, b = node_off[g0 + g + 1] - n0;
    double acc = 0.0;
    for (int i = a; i < b; ++i) acc += h[(long)i * F + f];
    const float inv_cnt = 1.f / (float)(b - a);
    const float pool = (float)acc * inv_cnt;
    const double dot = wave_sum((double)pool * w[GL::W_S + f]) + w[GL::C_S];
    const float s = (float)(1.0 / (1.0 + std::exp(-dot)));
    if (f == 0) sout[g] = s;
    if (!rows) return;
    const float t = ts[g0 + g];
    const float dp = 2.f * (s - t) * inv_b * s * (1.f - s);
    rows[(long)g * (F + 1) + f] = dp * pool;
    if (f == 0) {
        rows[(long)g * (F + 1) + F] = dp;
        loss[g] = (s - t) * (s - t) * inv_b;
    }
    const float d = dp * w[GL::W_S + f] * inv_cnt;
    for (int i = a; i < b; ++i) dh[(long)i * F + f] += d;
}

// dZ = dH_l+1 * [Z > 0], FP32 (bias gradient) and BF16 (MFMA operand).
__global__ void k_update_bwd(int nodes, const float* __restrict__ gh, const float* __restrict__ z,
                             float* __restrict__ dzf, uint16_t* __restrict__ dzb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes * F) return;
    const float d = z[idx] > 0.f ? gh[idx] : 0.f;
    dzf[idx] = d;
    dzb[idx] = to_bf16(d);
}

// Message backward at node x (lane = feature), with dM = dX[:, 64:128]:
//   as destination (y -> x): da = dM_x [P_y + Q_x + J w_J + b_m > 0] gives
//     dQ_x, and the rows [sum da J, sum da] of the w_J and b_m gradients;
//   as source (x -> y): dP_x = sum_y dM_y [P_x + Q_y + J w_J + b_m > 0].
__global__ __launch_bounds__(256) void k_msg_bwd(int nodes, int n0, const int32_t* __restrict__ adj_off,
                                                 const int32_t* __restrict__ adj_nbr, const float* __restrict__ adj_j,
                                                 const float* __restrict__ pq, const float* __restrict__ wl,
                                                 const float* __restrict__ dx, uint16_t* __restrict__ dpq,
                                                 float* __restrict__ rows) {
    const int x = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
    if (x >= nodes) return;
    const float px = pq[(long)x * 2 * F + f], qx = pq[(long)x * 2 * F + F + f];
    const float dmx = dx[(long)x * 2 * F + F + f];
    const float wj = wl[GL::W_J + f], bm = wl[GL::B_M + f];
    float dp = 0.f, dq = 0.f, sj = 0.f, sb = 0.f;
    const int e1 = adj_off[n0 + x + 1];
    for (int e = adj_off[n0 + x]; e < e1; ++e) {
        const int y = adj_nbr[e] - n0;
        const float jw = adj_j[e] * wj + bm;
        if (pq[(long)y * 2 * F + f] + qx + jw > 0.f) {
            dq += dmx;
            sj += dmx * adj_j[e];
            sb += dmx;
        }
        if (px + pq[(long)y * 2 * F + F + f] + jw > 0.f) dp += dx[(long)y * 2 * F + F + f];
    }
    dpq[(long)x * 2 * F + f] = to_bf16(dp);
    dpq[(long)x * 2 * F + F + f] = to_bf16(dq);
    rows[(long)x * 2 * F + f] = sj;
    rows[(long)x * 2 * F + F + f] = sb;
}

// dH_l = dH_l+1 (residual) + dX[:, 0:64] + dPQ W_sd^T.
__global__ void k_add_grad(int nodes, float* __restrict__ gh, const float* __restrict__ dx,
                           const float* __restrict__ dh2) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes * F) return;
    const int v = idx / F, f = idx % F;
    gh[idx] += dx[(long)v * 2 * F + f] + dh2[idx];
}

// Rows [x0 d, x1 d, d] with d = dH0 [pre > 0]: the W_in and b_in gradient rows.
__global__ void k_embed_bwd(int nodes, const float* __restrict__ feat, const float* __restrict__ pre,
                            const float* __restrict__ gh, float* __restrict__ rows) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes * F) return;
    const int v = idx / F, f = idx % F;
    const float d = pre[idx] > 0.f ? gh[idx] : 0.f;
    rows[(long)v * 3 * F + f] = feat[2 * v] * d;
    rows[(long)v * 3 * F + F + f] = feat[2 * v + 1] * d;
    rows[(long)v * 3 * F + 2 * F + f] = d;
}

// Adam on the FP32 master weights, then the BF16 operand copy.
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    wb[i] = to_bf16(wi);
}

__global__ void k_to_bf16(int n, const float* __restrict__ w, uint16_t* __restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = to_bf16(w[i]);
}

inline unsigned blocks_for(long work, int per) { return (unsigned)((work + per - 1) / per); }

constexpr int kChunk = 256;   // K rows per split of a weight-gradient product
constexpr int kMaxSplits = 1024;

}  // namespace

Gnn::Gnn(uint64_t seed) {
    TQ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from SplitMix64, zero biases.
    std::vector<float> w(GL::COUNT, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in) {
        const double a = 1.0 / std::sqrt((double)fan_in);
        for (int i = 0; i < count; ++i) w[off + i] = (float)((2.0 * next() - 1.0) * a);
    };
    fill(GL::W_IN, 2 * F, 2);
    for (int l = 0; l < GL::L; ++l) {
        const int b = GL::layer_base(l);
        fill(b + GL::W_SD, 2 * F * F, F);
        fill(b + GL::W_J, F, 1);
        fill(b + GL::W_U, 2 * F * F, 2 * F);
    }
    fill(GL::W_Z, F, F);
    fill(GL::W_S, F, F);
    w_.alloc(GL::COUNT);
    g_.alloc(GL::COUNT);
    m_.alloc(GL::COUNT);
    v_.alloc(GL::COUNT);
    wb_.alloc(GL::COUNT);
    lossv_.alloc(2);
    set_params(w.data());
}

Gnn::~Gnn() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Gnn::get_params(float* out) {
    TQ_HIP(hipMemcpyAsync(out, w_.p, GL::COUNT * sizeof(float), hipMemcpyDeviceToHost, stream_));
    TQ_HIP(hipStreamSynchronize(stream_));
}

void Gnn::set_params(const float* in) {
    for (int i = 0; i < GL::COUNT; ++i)
        if (!std::isfinite(in[i])) throw Error(-4, "non-finite parameter");
    TQ_HIP(hipMemcpyAsync(w_.p, in, GL::COUNT * sizeof(float), hipMemcpyHostToDevice, stream_));
    std::vector<float> zero(GL::COUNT, 0.f);
    TQ_HIP(hipMemcpyAsync(m_.p, zero.data(), GL::COUNT * sizeof(float), hipMemcpyHostToDevice, stream_));
    TQ_HIP(hipMemcpyAsync(v_.p, zero.data(), GL::COUNT * sizeof(float), hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(GL::COUNT, 256)), dim3(256), 0, stream_, GL::COUNT, w_.p, wb_.p);
    TQ_LAUNCH_CHECK();
    TQ_HIP(hipStreamSynchronize(stream_));
    adam_t_ = 0;
}

void Gnn::set_data(int graphs, const int32_t* node_off, const float* feat, const int32_t* adj_off,
                   const int32_t* adj_nbr, const float* adj_j, const float* z_target, const float* s_target) {
    if (graphs <= 0) throw Error(-1, "need at least one graph");
    if (node_off[0] != 0) throw Error(-1, "node_off[0] must be 0");
    for (int g = 0; g < graphs; ++g)
        if (node_off[g + 1] <= node_off[g]) throw Error(-1, "every graph needs at least one node");
    const int nodes = node_off[graphs];
    // per-node kernels index nodes * 192 elements in 32-bit arithmetic
    if (nodes > (1 << 24)) throw Error(-1, "more than 2^24 nodes in one dataset");
    if (adj_off[0] != 0) throw Error(-1, "adj_off[0] must be 0");
    for (int i = 0; i < nodes; ++i)
        if (adj_off[i + 1] < adj_off[i]) throw Error(-1, "adj_off must be non-decreasing");
    const int nnz = adj_off[nodes];
    for (int g = 0; g < graphs; ++g)
        for (int i = node_off[g]; i < node_off[g + 1]; ++i)
            for (int e = adj_off[i]; e < adj_off[i + 1]; ++e)
                if (adj_nbr[e] < node_off[g] || adj_nbr[e] >= node_off[g + 1])
                    throw Error(-1, "neighbour outside its graph");
    for (long i = 0; i < 2L * nodes; ++i)
        if (!std::isfinite(feat[i])) throw Error(-4, "non-finite node feature");
    for (int e = 0; e < nnz; ++e)
        if (!std::isfinite(adj_j[e])) throw Error(-4, "non-finite coupling");
    const bool targets = z_target && s_target;
    if (targets) {
        for (int i = 0; i < nodes; ++i)
            if (!std::isfinite(z_target[i])) throw Error(-4, "non-finite target");
        for (int g = 0; g < graphs; ++g)
            if (!std::isfinite(s_target[g])) throw Error(-4, "non-finite target");
    }
    // everything is checked before the previous dataset is replaced
    TQ_HIP(hipStreamSynchronize(stream_));
    graphs_ = graphs;
    nodes_ = nodes;
    node_off_h_.assign(node_off, node_off + graphs + 1);
    adj_off_h_.assign(adj_off, adj_off + nodes + 1);
    node_off_.alloc(graphs + 1);
    adj_off_.alloc(nodes + 1);
    adj_nbr_.alloc(std::max(nnz, 1));
    feat_.alloc(2 * (std::size_t)nodes);
    adj_j_.alloc(std::max(nnz, 1));
    TQ_HIP(hipMemcpy(node_off_.p, node_off, (graphs + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    TQ_HIP(hipMemcpy(adj_off_.p, adj_off, (nodes + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    if (nnz) {
        TQ_HIP(hipMemcpy(adj_nbr_.p, adj_nbr, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(adj_j_.p, adj_j, nnz * sizeof(float), hipMemcpyHostToDevice));
    }
    TQ_HIP(hipMemcpy(feat_.p, feat, 2 * (std::size_t)nodes * sizeof(float), hipMemcpyHostToDevice));
    has_targets_ = targets;
    if (has_targets_) {
        tz_.alloc(nodes);
        ts_.alloc(graphs);
        TQ_HIP(hipMemcpy(tz_.p, z_target, nodes * sizeof(float), hipMemcpyHostToDevice));
        TQ_HIP(hipMemcpy(ts_.p, s_target, graphs * sizeof(float), hipMemcpyHostToDevice));
    }
}

void Gnn::ensure(int nodes, int gr) {
    if (nodes <= cap_nodes_ && gr <= cap_graphs_) return;
    TQ_HIP(hipStreamSynchronize(stream_));
    const std::size_t N = (std::size_t)std::max(nodes, cap_nodes_);
    const std::size_t B = (std::size_t)std::max(gr, cap_graphs_);
    pre_in_.alloc(N * F);
    for (int l = 0; l <= GL::L; ++l) hf_[l].alloc(N * F);
    for (int l = 0; l < GL::L; ++l) {
        pq_[l].alloc(N * 2 * F);
        zs_[l].alloc(N * F);
        xb_[l].alloc(N * 2 * F);
    }
    z_.alloc(N);
    s_.alloc(B);
    gh_.alloc(N * F);
    dzf_.alloc(N * F);
    dzb_.alloc(N * F);
    dxf_.alloc(N * 2 * F);
    dpqb_.alloc(N * 2 * F);
    dh2_.alloc(N * F);
    rows_.alloc(std::max(N * 3 * F, B * (F + 1)));
    lossrows_.alloc(std::max(N, B));
    const std::size_t splits = std::min<std::size_t>(kMaxSplits, (N + kChunk - 1) / kChunk);
    part_.alloc(splits * 2 * F * F);
    red_.alloc(256 * (std::size_t)std::max<std::size_t>(3 * F, 2 * F * F));
    cap_nodes_ = (int)N;
    cap_graphs_ = (int)B;
}

void Gnn::gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, int M, int N,
               int K, bool wgrad) {
    int kchunk = K, splits = 1;
    if (wgrad) {
        kchunk = std::max(kChunk, ((K + kMaxSplits - 1) / kMaxSplits + 15) / 16 * 16);
        splits = (K + kchunk - 1) / kchunk;
        if ((std::size_t)splits * M * N > part_.n) throw Error(-1, "gnn: split buffer too small");
    }
    float* dst = wgrad ? part_.p : C;
    hipLaunchKernelGGL(k_gemm_bf16, dim3(blocks_for(N, 32), blocks_for(M, 32), splits), dim3(256), 0, stream_, A, sam,
                       sak, B, sbk, sbn, dst, M, N, K, kchunk);
    TQ_LAUNCH_CHECK();
    if (wgrad) colsum(part_.p, splits, M * N, C);
}

void Gnn::colsum(const float* X, int R, int C, float* out) {
    const int nblk = std::max(1, std::min(256, (R + 63) / 64));
    const int rows = (R + nblk - 1) / nblk;
    if (nblk == 1) {
        hipLaunchKernelGGL(k_colsum, dim3(blocks_for(C, 64), 1), dim3(256), 0, stream_, X, R, C, std::max(rows, 1),
                           out);
        TQ_LAUNCH_CHECK();
        return;
    }
    if ((std::size_t)nblk * C > red_.n) throw Error(-1, "gnn: reduction buffer too small");
    hipLaunchKernelGGL(k_colsum, dim3(blocks_for(C, 64), nblk), dim3(256), 0, stream_, X, R, C, rows, red_.p);
    hipLaunchKernelGGL(k_colsum, dim3(blocks_for(C, 64), 1), dim3(256), 0, stream_, (const float*)red_.p, nblk, C,
                       nblk, out);
    TQ_LAUNCH_CHECK();
}

void Gnn::forward(int g0, int g1, bool train) {
    if (g0 < 0 || g1 > graphs_ || g0 >= g1) throw Error(-1, "gnn: bad graph range");
    if (train && !has_targets_) throw Error(-1, "gnn: no targets set");
    const int n0 = node_off_h_[g0];
    const int nodes = node_off_h_[g1] - n0;
    const int ng = g1 - g0;
    ensure(nodes, ng);
    const unsigned eb = blocks_for((long)nodes * F, 256);
    const unsigned wb4 = blocks_for(nodes, 4);
    hipLaunchKernelGGL(k_embed_fwd, dim3(eb), dim3(256), 0, stream_, nodes, (const float*)feat_.p + 2L * n0,
                       (const float*)w_.p, pre_in_.p, hf_[0].p, xb_[0].p);
    for (int l = 0; l < GL::L; ++l) {
        const int base = GL::layer_base(l);
        // [P Q] = H_l W_sd: A = X_l[:, 0:64] (row stride 128), B = W_sd (64 x 128)
        gemm(xb_[l].p, 2 * F, 1, wb_.p + base + GL::W_SD, 2 * F, 1, pq_[l].p, nodes, 2 * F, F, false);
        hipLaunchKernelGGL(k_msg_fwd, dim3(wb4), dim3(256), 0, stream_, nodes, n0, (const int32_t*)adj_off_.p,
                           (const int32_t*)adj_nbr_.p, (const float*)adj_j_.p, (const float*)pq_[l].p,
                           (const float*)w_.p + base, xb_[l].p);
        // Z = X_l W_u: A = X_l (N x 128), B = W_u (128 x 64)
        gemm(xb_[l].p, 2 * F, 1, wb_.p + base + GL::W_U, F, 1, zs_[l].p, nodes, F, 2 * F, false);
        hipLaunchKernelGGL(k_update_fwd, dim3(eb), dim3(256), 0, stream_, nodes, zs_[l].p,
                           (const float*)w_.p + base, (const float*)hf_[l].p, hf_[l + 1].p,
                           l + 1 < GL::L ? xb_[l + 1].p : (uint16_t*)nullptr);
    }
    TQ_LAUNCH_CHECK();
    float* rows = train ? rows_.p : nullptr;
    hipLaunchKernelGGL(k_head_node, dim3(wb4), dim3(256), 0, stream_, nodes, n0, (const float*)hf_[GL::L].p,
                       (const float*)w_.p, (const float*)tz_.p, z_.p, rows, gh_.p, lossrows_.p, 1.f / nodes);
    TQ_LAUNCH_CHECK();
    if (train) {  // node-head gradient rows and loss are reduced before the graph head reuses the buffers
        colsum(rows_.p, nodes, F + 1, g_.p + GL::W_Z);
        colsum(lossrows_.p, nodes, 1, lossv_.p);
    }
    hipLaunchKernelGGL(k_head_graph, dim3(blocks_for(ng, 4)), dim3(256), 0, stream_, ng, g0,
                       (const int32_t*)node_off_.p, (const float*)hf_[GL::L].p, (const float*)w_.p,
                       (const float*)ts_.p, s_.p, rows, gh_.p, lossrows_.p, 1.f / ng);
    TQ_LAUNCH_CHECK();
    if (train) {
        colsum(rows_.p, ng, F + 1, g_.p + GL::W_S);
        colsum(lossrows_.p, ng, 1, lossv_.p + 1);
    }
}

double Gnn::backward(int g0, int g1) {
    const int n0 = node_off_h_[g0];
    const int nodes = node_off_h_[g1] - n0;
    const unsigned eb = blocks_for((long)nodes * F, 256);
    const unsigned wb4 = blocks_for(nodes, 4);
    for (int l = GL::L - 1; l >= 0; --l) {
        const int base = GL::layer_base(l);
        hipLaunchKernelGGL(k_update_bwd, dim3(eb), dim3(256), 0, stream_, nodes, (const float*)gh_.p,
                           (const float*)zs_[l].p, dzf_.p, dzb_.p);
        TQ_LAUNCH_CHECK();
        colsum(dzf_.p, nodes, F, g_.p + base + GL::B_U);
        // dW_u = X_l^T dZ (128 x 64), K = nodes
        gemm(xb_[l].p, 1, 2 * F, dzb_.p, F, 1, g_.p + base + GL::W_U, 2 * F, F, nodes, true);
        // dX = dZ W_u^T (N x 128)
        gemm(dzb_.p, F, 1, wb_.p + base + GL::W_U, 1, F, dxf_.p, nodes, 2 * F, F, false);
        hipLaunchKernelGGL(k_msg_bwd, dim3(wb4), dim3(256), 0, stream_, nodes, n0, (const int32_t*)adj_off_.p,
                           (const int32_t*)adj_nbr_.p, (const float*)adj_j_.p, (const float*)pq_[l].p,
                           (const float*)w_.p + base, (const float*)dxf_.p, dpqb_.p, rows_.p);
        TQ_LAUNCH_CHECK();
        colsum(rows_.p, nodes, 2 * F, g_.p + base + GL::W_J);  // [w_J | b_m] are adjacent
        // dW_sd = H_l^T dPQ (64 x 128): A(m,k) = X_l[k, m]
        gemm(xb_[l].p, 1, 2 * F, dpqb_.p, 2 * F, 1, g_.p + base + GL::W_SD, F, 2 * F, nodes, true);
        // dPQ W_sd^T (N x 64)
        gemm(dpqb_.p, 2 * F, 1, wb_.p + base + GL::W_SD, 1, 2 * F, dh2_.p, nodes, F, 2 * F, false);
        hipLaunchKernelGGL(k_add_grad, dim3(eb), dim3(256), 0, stream_, nodes, gh_.p, (const float*)dxf_.p,
                           (const float*)dh2_.p);
        TQ_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_embed_bwd, dim3(eb), dim3(256), 0, stream_, nodes, (const float*)feat_.p + 2L * n0,
                       (const float*)pre_in_.p, (const float*)gh_.p, rows_.p);
    TQ_LAUNCH_CHECK();
    colsum(rows_.p, nodes, 3 * F, g_.p + GL::W_IN);  // [W_in row 0 | W_in row 1 | b_in]
    float loss[2];
    TQ_HIP(hipMemcpyAsync(loss, lossv_.p, 2 * sizeof(float), hipMemcpyDeviceToHost, stream_));
    TQ_HIP(hipStreamSynchronize(stream_));
    return (double)loss[0] + (double)loss[1];
}

void Gnn::predict(int g0, int g1, float* out_z, float* out_s) {
    forward(g0, g1, false);
    const int n0 = node_off_h_[g0];
    const int nodes = node_off_h_[g1] - n0;
    if (out_z) TQ_HIP(hipMemcpyAsync(out_z, z_.p, nodes * sizeof(float), hipMemcpyDeviceToHost, stream_));
    if (out_s) TQ_HIP(hipMemcpyAsync(out_s, s_.p, (g1 - g0) * sizeof(float), hipMemcpyDeviceToHost, stream_));
    TQ_HIP(hipStreamSynchronize(stream_));
}

double Gnn::grad(int g0, int g1, float* out_grad) {
    forward(g0, g1, true);
    const double loss = backward(g0, g1);
    if (out_grad) {
        TQ_HIP(hipMemcpyAsync(out_grad, g_.p, GL::COUNT * sizeof(float), hipMemcpyDeviceToHost, stream_));
        TQ_HIP(hipStreamSynchronize(stream_));
    }
    return loss;
}

double Gnn::train_step(int g0, int g1, float lr) {
    if (!(lr > 0.f) || !std::isfinite(lr)) throw Error(-1, "gnn: learning rate must be positive");
    forward(g0, g1, true);
    const double loss = backward(g0, g1);
    ++adam_t_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)adam_t_), bc2 = 1.f - std::pow(b2, (float)adam_t_);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(GL::COUNT, 256)), dim3(256), 0, stream_, GL::COUNT, w_.p,
                       (const float*)g_.p, m_.p, v_.p, wb_.p, lr, b1, b2, 1e-8f, bc1, bc2);
    TQ_LAUNCH_CHECK();
    return loss;
}

}  // namespace tq
