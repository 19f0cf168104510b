// MPS-TEBD engine on gfx950 (upstream cuda/src/mps.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "common.hpp"

namespace tq {

struct TebdModel {
    int n = 0;
    std::vector<double> h;
    std::vector<int> ei, ej;  // as given; routing normalises i < j (mps.cpp:127)
    std::vector<double> ew;
    std::vector<double> bias;  // empty or n entries
};

// Diagnostics of the SVD kernels: calls, Jacobi sweeps, column-pair visits and
// rotations weighted by the column length (pair-rows), and the HIP-event time of
// the SVD launches while timing is on.
struct SvdStats {
    long long calls = 0, sweeps = 0;
    double visited_rows = 0, rotated_rows = 0;
    double svd_ms = 0;
    long long timed_launches = 0;
    long long supersteps = 0, exec_updates = 0;  // lockstep executor: super-steps run, updates placed
};

// A Trotter step's values on the device, for a step captured once and replayed
// (hipGraph): schedule values (a[k], b[k]) at ab[2k], ZZ gate phasors (cos, sin
// of -b[k] J_e dt / ns) at zz[k * m + e], par = {k, first recorded entropy slot}.
// par[2] = first max-bond slot. All null outside a capture: kernels then take
// their launch arguments.
struct StepParams {
    const double* ab = nullptr;
    const double2* zz = nullptr;
    const int* par = nullptr;
    int m = 0;
};

// One two-site update of the routed ZZ block.
struct Update {
    int site;   // acts on (site, site+1)
    int edge;   // -1 for SWAP, else index of the ZZ edge
};

class Tebd {
public:
    Tebd(const TebdModel& model, int chi, double dt, int zz_substeps);
    ~Tebd();
    Tebd(const Tebd&) = delete;
    Tebd& operator=(const Tebd&) = delete;

    // |+>^n, or with `warm` the product of the single-spin ground states of
    // -(X_i + warm_i Z_i) (the s = 0 ground state under a warm-start bias).
    void reset_product(const double* warm = nullptr);
    // Random MPS with every bond at capacity, then one left-to-right sweep of
    // identity updates giving bond weights ~exp(-k/decay) (a saturated state).
    void randomize(uint64_t seed, double decay);
    // Enqueue one Trotter step (mps.cpp:309-356); entropy recorded at slot `rec` if >= 0.
    void step(double A, double B, int rec);
    // Enqueue `count` steps with schedule values a[k], b[k]; when rec0 >= 0 the
    // entropy of step k is recorded at slot rec0 + k. Runs the steps one launch
    // sequence per update, or as one persistent launch when the executor is on.
    // bond0 >= 0: the largest bond after step k is kept at max-bond slot bond0 + k
    // (read with max_bonds after sync; the executor takes it from the entropy
    // snapshots, so without recording the steps run on the per-update path).
    void run(const double* a, const double* b, int count, int rec0, int bond0 = -1);
    // Lockstep executor: 0 = off, g > 0 = at most g groups of workgroups,
    // -1 = as many groups as are co-resident (capped at 64). workers: workgroups
    // per group (0 = one per block pair of a tournament round).
    void set_executor(int groups, int workers = 0);
    int executor_groups() const { return exec_req_; }
    int max_capacity() const { return maxcap_; }  // largest bond capacity min(chi, 2^b, 2^(n-b))
    int executor_workers() const { return exec_workers_; }
    // LQ preconditioning of the multi-workgroup Jacobi SVD (on by default); in a
    // batched run the first session's setting applies.
    void set_svd_lq(bool on);
    bool svd_lq() const { return svd_lq_; }
    // Batched lockstep run: instance i advances counts[i] steps with schedule
    // values as[i][k], bs[i][k]; entropy recording (rec0 >= 0) needs one instance.
    // mode: 0 auto (solo for four or more instances with chi <= 32), 1 grouped
    // (an update's SVD spread over a group of workgroups), 2 solo (one
    // workgroup per update).
    // bond0 >= 0 (with rec0 >= 0): the largest bond after each recorded step into
    // max-bond slot bond0 + k, from the snapshots' bond dimensions.
    static void run_lockstep(const std::vector<Tebd*>& ts, const std::vector<const double*>& as,
                             const std::vector<const double*>& bs, const std::vector<int>& counts, int rec0, int req,
                             int mode = 0, int bond0 = -1);
    void sync();
    std::vector<double> entropy(int count);      // recorded slots [0, count)
    double entropy_now();                        // qubit-0 entropy of the current state (mps.cpp:218-291)
    // Per-spin longitudinal field added to the single-site generator from the
    // next step on: -A X - (B h_i + A bias_i) Z; nullptr removes it.
    void set_bias(const double* bias);
    double max_trunc();
    SvdStats svd_stats();
    // Executor phase profile of workgroup 0 since creation, in ms (slots 0-6,
    // 8, 10) or counts (7 updates run by its group, 9 rounds, 11 sweeps).
    std::vector<double> exec_profile();
    void set_timing(bool on);  // HIP events around every SVD launch
    std::vector<int> bonds();                    // n+1 bond dims
    struct Readout {
        std::vector<int8_t> spins;
        double energy = 0, expected_energy = 0, prob = 0;
        std::vector<double> z;  // <Z_i> / <psi|psi>
    };
    Readout readout();                           // mps.cpp:358-380
    std::vector<double> dense();                 // n <= 24, interleaved re/im
    std::vector<double> theta(int p, int& rows, int& cols);  // identity-gate theta at (p, p+1)
    std::size_t updates_per_step() const { return plan_.size() * (std::size_t)ns_; }
    // Device time (ms) from the start of the last executor launch that advanced
    // this session to the completion of its last update: the end of the super-step
    // holding it, or in the dataflow order its latest done flag (0 when that launch
    // gave it no update).
    double last_run_ms() const { return last_run_ms_; }
    int n() const { return n_; }
    hipStream_t stream() const { return stream_; }
    // Capacity of bond b (0..n): min(chi, 2^b, 2^(n-b)).
    int cap(int b) const { return cap_[b]; }
    void ensure_trace(int count);
    // Largest bond after each step, kept on the device: size the buffer for
    // `count` steps, record slot k after step k (enqueued, no host sync), read
    // slots [0, count) after sync().
    void ensure_max_bond(int count);
    void record_max_bond(int slot);
    std::vector<int> max_bonds(int count);

private:
    void build_plan();
    void run_exec(const double* a, const double* b, int count, int rec0, int bond0 = -1);
    void two_site(int p, const Gate4& g, const double* wts = nullptr, int edge = -1);
    void run_graph(const double* a, const double* b, int count, int rec0, int bond0);
    void drop_graph();
    void entropy_sweep(int rec);
    void entropy_of(const double2* sites, const int* dims, int rec);
    void entropy_into(const double2* sites, const int* dims, double* dst);
    void env_left(const double2* lin, int site, bool z, double2* lout);
    void env_right(const double2* rin, int site, double2* rout, const double2* sites = nullptr,
                   const int* dims = nullptr);
    double2* site_ptr(int i) const { return sites_.p + off_[i]; }
    int chi_max() const { return maxcap_; }

    int n_, chi_, ns_;
    double dt_;
    TebdModel model_;
    std::vector<int> cap_;
    std::vector<std::size_t> off_;
    std::vector<Update> plan_;  // one ZZ slice in forward order
    std::vector<std::pair<std::size_t, std::size_t>> spans_;  // plan_ range of each edge
    int maxcap_ = 1;
    hipStream_t stream_ = nullptr;

    DevBuf<double2> sites_;
    DevBuf<int> dims_;
    DevBuf<int> caps_dev_;
    DevBuf<std::size_t> off_dev_;
    DevBuf<double2> mat_, work_;   // theta (kept) and the Jacobi working copy
    DevBuf<double2> aux_, taus_;   // LQ reflectors of the distributed SVD
    DevBuf<double> sv_;            // sorted singular values
    DevBuf<int> perm_;             // sorted position -> column
    DevBuf<double> trunc_;         // running max discarded weight
    DevBuf<int> flags_;            // convergence flags of the distributed SVD
    DevBuf<unsigned long long> stats_;  // calls, sweeps, visited pair-rows, rotated pair-rows
    bool timing_ = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    std::size_t ev_next_ = 0;
    double timed_ms_ = 0;
    long long timed_n_ = 0;
    int ld_ = 0;                   // leading dimension of mat_/work_
    bool force_dist_ = false;
    bool small_svd_ = true;  // thetas up to 64 columns x 128 rows take k_svd_small
    int xcd_stride_ = 1;
    DevBuf<double> h_, bias_;
    DevBuf<double2> envA_, envB_, envT_;
    DevBuf<double> trace_;
    int trace_cap_ = 0;
    DevBuf<int> mbond_;
    int mbond_cap_ = 0;
    std::size_t svd_lds_ = 0;

    // lockstep executor state
    bool svd_lq_ = true;
    int exec_req_ = 0;
    int exec_workers_ = 0;
    long long exec_supersteps_ = 0, exec_updates_ = 0;
    double last_run_ms_ = 0.0;             // see last_run_ms()
    DevBuf<unsigned long long> ex_clock_;  // super-step end clocks of the last launch
    DevBuf<unsigned char> ex_tasks_;   // schedule: updates, snapshots, rounds, snapshot offsets
    DevBuf<double2> ex_mat_, ex_work_, ex_snap_, ex_gates_, ex_aux_, ex_taus_;
    DevBuf<double> ex_sv_;
    DevBuf<int> ex_perm_, ex_flags_, ex_sdims_;
    DevBuf<int> ex_bar_;                   // the barrier error word, then group barriers [G][2]
    DevBuf<unsigned long long> ex_prof_;   // phase profile of workgroup 0
    bool ex_group_sync_ = true;            // grouped layout: per-group barriers inside an update
    bool ex_xcd_local_ = false;            // ... XCD-local when the groups sit one per XCD (opt-in)
    bool ex_armed_ = false;                // a launch whose barrier error word is still unread
    long long ex_spin_limit_ = 1ll << 24;  // GroupSync polls before a wait gives up
    int ex_naps_ = 1;                      // 64-cycle sleeps between the polls of a wait
    double svd_stop_ = 0.0;                // SvdMode::stop_ratio of the executor's SVDs
    bool svd_rows_ = true;                 // SvdMode::rows of the executor's SVDs
    bool svd_gram_ = true;                 // SvdMode::gram of the executor's SVDs
    bool ex_step_prof_ = false;            // cycle profile of the row-layout steps
    bool ex_prof_on_ = true;               // workgroup 0's phase profile
    bool ex_p2p_ = true;                   // per-group layout: point-to-point tournament rounds
    DevBuf<int> ex_ep_;                    // their round counters
    int ex_dataflow_ = 1;   // per-group layout: updates wait for their predecessors only (0: super-step barriers)
    bool ex_queue_ = true;  // dataflow order: groups take the next update from a shared queue
    DevBuf<int> ex_done_;                  // their done flags, then the queue's counter and claims
    std::vector<unsigned char> ex_host_;

    // per-update path captured as a hipGraph (run_graph)
    bool graph_on_ = true;         // knob TESSERA_STEP_GRAPH=0 turns it off
    bool plan_graphable_ = false;  // no update needs the cooperative SVD launch
    bool capturing_ = false;
    StepParams sp_;                // device step values while capturing
    DevBuf<double> g_ab_;
    DevBuf<double2> g_zz_;
    DevBuf<int> g_par_;
    std::vector<double> g_host_ab_;
    std::vector<double2> g_host_zz_;
    int g_host_par_[3] = {0, 0, 0};
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    int graph_rec_ = -1;
    const void* graph_key_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

}  // namespace tq
