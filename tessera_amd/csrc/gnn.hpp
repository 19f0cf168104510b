// Message-passing network for learned anneal guidance on gfx950: BF16 operands
// on MFMA with FP32 accumulation, FP32 master weights and Adam moments.
//
// No upstream kernel counterpart. The upstream python/gnn_guide.py:81-103
// (IsingGNN: node embedding, edge-conditioned messages summed at the
// destination, node update, per-node readout) is the structural template; the
// layer algebra, targets and heads here are this library's own:
//
//   x_i   = (h_i, deg_i / 3)                              node features
//   H0    = relu(x W_in + b_in)                           width F = 64
//   layer l = 0..2:
//     [P Q] = H_l W_sd                                    (N x 64)(64 x 128) on MFMA
//     a_uv  = relu(P_u + Q_v + J_uv w_J + b_m)            message u -> v
//     M_v   = sum_{u in adj(v)} a_uv
//     Z     = [H_l M] W_u + b_u                           (N x 128)(128 x 64) on MFMA
//     H_l+1 = H_l + relu(Z)
//   z_i   = tanh(H3_i . w_z + c_z)                        per-spin <Z_i> estimate
//   s_g   = sigmoid(mean_{i in g} H3_i . w_s + c_s)       per-graph entanglement-peak time
//   loss  = mean_i (z_i - tz_i)^2 + mean_g (s_g - ts_g)^2
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace tq {

struct GnnLayout {
    static constexpr int F = 64, L = 3;
    static constexpr int W_IN = 0, B_IN = 2 * F;
    static constexpr int LAYER0 = B_IN + F;
    // per-layer offsets relative to layer_base(l)
    static constexpr int W_SD = 0, W_J = F * 2 * F, B_M = W_J + F, W_U = B_M + F, B_U = W_U + 2 * F * F;
    static constexpr int LAYER_SIZE = B_U + F;
    static constexpr int HEAD = LAYER0 + L * LAYER_SIZE;
    static constexpr int W_Z = HEAD, C_Z = W_Z + F, W_S = C_Z + 1, C_S = W_S + F;
    static constexpr int COUNT = C_S + 1;
    static constexpr int layer_base(int l) { return LAYER0 + l * LAYER_SIZE; }
};

class Gnn {
public:
    explicit Gnn(uint64_t seed);
    ~Gnn();
    Gnn(const Gnn&) = delete;
    Gnn& operator=(const Gnn&) = delete;

    void get_params(float* out);
    void set_params(const float* in);  // also resets the Adam moments
    // Device-resident dataset of `graphs` graphs; nodes of graph g are
    // [node_off[g], node_off[g+1]); adjacency of node i is
    // adj_off[i]..adj_off[i+1] (neighbour ids are global node ids, symmetric).
    void set_data(int graphs, const int32_t* node_off, const float* feat, const int32_t* adj_off,
                  const int32_t* adj_nbr, const float* adj_j, const float* z_target, const float* s_target);
    void predict(int g0, int g1, float* out_z, float* out_s);
    double grad(int g0, int g1, float* out_grad);  // loss; gradient of the batch mean loss
    double train_step(int g0, int g1, float lr);   // one Adam step, returns the loss before it
    int graphs() const { return graphs_; }

private:
    void forward(int g0, int g1, bool train);
    double backward(int g0, int g1);
    // C = A B with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (BF16),
    // FP32 result M x N row-major; with wgrad the K range is split over
    // workgroups and the partial products are summed by a second pass.
    void gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, int M, int N,
              int K, bool wgrad);
    void colsum(const float* X, int R, int C, float* out);
    void ensure(int nodes, int gr);

    hipStream_t stream_ = nullptr;
    int graphs_ = 0, nodes_ = 0;
    bool has_targets_ = false;
    std::vector<int32_t> node_off_h_, adj_off_h_;
    // parameters (FP32 master), BF16 copies, gradient, Adam moments
    DevBuf<float> w_, g_, m_, v_;
    DevBuf<uint16_t> wb_;
    long long adam_t_ = 0;
    // dataset
    DevBuf<int32_t> node_off_, adj_off_, adj_nbr_;
    DevBuf<float> feat_, adj_j_, tz_, ts_;
    // activations (per batch, grown on demand)
    int cap_nodes_ = 0, cap_graphs_ = 0;
    DevBuf<float> pre_in_, hf_[GnnLayout::L + 1], pq_[GnnLayout::L], zs_[GnnLayout::L];
    DevBuf<uint16_t> xb_[GnnLayout::L];
    DevBuf<float> z_, s_, pool_;
    // backward work
    DevBuf<float> gh_, dzf_, dxf_, dh2_, rows_, part_, lossrows_, red_, lossv_;
    DevBuf<uint16_t> dzb_, dpqb_;
};

}  // namespace tq
