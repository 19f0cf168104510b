/*
 * tessera_qa: MI355X (gfx950) implementation of TESSERA's quantum-annealing hot
 * path behind the upstream C ABI.
 *
 *   H(s) = -A(s) sum_i X_i - B(s) ( sum_<ij> J_ij Z_i Z_j + sum_i h_i Z_i )
 *
 * Two engines:
 *   - MPS-TEBD (upstream cuda/src/mps.cpp): second-order Trotter, single-site
 *     half steps around a block of ZZ gates routed by swap networks, every
 *     two-site gate followed by an SVD truncated to chi.
 *   - exact state vector (upstream rust/src/quantum.rs): phase(B dt/2),
 *     X-rotations(A dt), phase(B dt/2) on all 2^n amplitudes.
 *
 * Conventions shared with upstream:
 *   - spins are int8 in {-1,+1}; MPS physical index 0 <-> +1; state-vector bit
 *     set <-> +1;
 *   - energies sum edges in list order, then fields (rust/src/ising.rs:23-32);
 *   - schedules are caller-sampled arrays a[k], b[k], k < steps;
 *   - all buffers are caller-owned host memory, contiguous, and outlive the
 *     call; the library keeps no pointer after returning.
 *
 * Every entry point is synchronous and may be called from any host thread.
 * Device work runs on the first visible GPU.
 */
#ifndef TESSERA_QA_H
#define TESSERA_QA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the extended entry points. */
enum {
    TESSERA_OK = 0,
    TESSERA_ERR_ARG = -1,        /* null pointer, n == 0, steps < 2, bad edge index */
    TESSERA_ERR_TOO_LARGE = -2,  /* state does not fit the device */
    TESSERA_ERR_DEVICE = -3,     /* HIP runtime or kernel launch failure */
    TESSERA_ERR_NONFINITE = -4   /* NaN or Inf in h, J, schedule or dt */
};

/*
 * Drop-in for upstream cuda/src/ffi.cpp:10-16 (bound by rust/src/ffi.rs:19-35):
 * MPS-TEBD with dt = 0.1 and one ZZ slice per step (ffi.cpp:28-29). Writes
 * out_s[n] (always) and out_entropy[steps] (qubit-0 entropy after each step,
 * if non-null); returns the Ising energy of the readout spins
 * sign(<Z_i>), 0 -> +1. Edge indices are narrowed to int as upstream does.
 * Never throws; on failure writes out_s = +1 and returns NaN (upstream has no
 * error path; tessera_last_error() tells what failed).
 */
double tessera_qa_anneal(const double* h, size_t n,
                         const size_t* ei, const size_t* ej, const double* ew, size_t m,
                         const double* schedule_a, const double* schedule_b, size_t steps,
                         size_t chi, int8_t* out_s, double* out_entropy);

/*
 * Extended MPS-TEBD entry point (superset of tessera_qa_anneal; dt = 0.1,
 * zz_substeps = 1, bias = NULL reproduces it). Replaces the upstream
 * mps_anneal(...) signature (cuda/include/mps.hpp:76-82) and exposes the
 * AnnealResult fields the plain ABI drops (mps.hpp:37-44).
 *   dt            Trotter step
 *   zz_substeps   ZZ block slices, alternating forward/reverse (mps.cpp:330-341)
 *   bias          optional per-spin longitudinal warm-start field: the
 *                 single-site generator becomes -A X - (B h_i + A bias_i) Z,
 *                 which vanishes with A at the end of the anneal (NULL = none)
 *   out_scalars   [4] = energy, expected_energy <psi|H|psi>, prob of the readout,
 *                 max_trunc (largest discarded weight over the run)
 *   out_entropy   [steps] qubit-0 entropy trace (nullable)
 *   out_max_bond  [steps] largest bond after each step (nullable)
 * Returns TESSERA_OK or a negative status.
 */
int tessera_qa_anneal_ex(const double* h, size_t n,
                         const size_t* ei, const size_t* ej, const double* ew, size_t m,
                         const double* schedule_a, const double* schedule_b, size_t steps,
                         double dt, size_t chi, int zz_substeps, const double* bias,
                         int8_t* out_s, double* out_scalars, double* out_entropy,
                         int32_t* out_max_bond);

/*
 * Exact state-vector anneal with an arbitrary schedule: the C form of
 * rust/src/quantum.rs:63 anneal_statevector (QaOutcome fields, quantum.rs:27-43).
 * No n <= 28 cap: n is limited by device memory (16 * 2^n bytes).
 *   out_s[n]                 argmax configuration (first index on ties)
 *   out_prob                 success probability of out_s            (nullable)
 *   out_entropy[steps]       qubit-0 entropy trace                    (nullable)
 *   out_energy_trace[steps]  <H_problem> after each step              (nullable)
 *   out_energy               Ising energy of out_s                    (nullable)
 *   out_expected_energy      final <H_problem>                        (nullable)
 */
int tessera_sv_anneal(const double* h, size_t n,
                      const size_t* ei, const size_t* ej, const double* ew, size_t m,
                      const double* schedule_a, const double* schedule_b, size_t steps, double dt,
                      int8_t* out_s, double* out_prob, double* out_entropy, double* out_energy_trace,
                      double* out_energy, double* out_expected_energy);

/*
 * tessera_sv_anneal for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * upstream counterpart. Same arguments, results and errors as
 * tessera_sv_anneal, but every device buffer of the anneal (state, schedule
 * tables, traces, per-block partials) lives in one static 4 MiB array per GPU
 * held until the process ends; the call performs no hipMalloc/hipFree.
 * Limits: n <= 16, steps <= 16384, m <= 4096 (TESSERA_ERR_TOO_LARGE beyond).
 * Calls on one GPU share the array: they are serialised inside a process, and
 * each call overwrites what the previous one left there.
 */
int tessera_sv_anneal_static(const double* h, size_t n,
                             const size_t* ei, const size_t* ej, const double* ew, size_t m,
                             const double* schedule_a, const double* schedule_b, size_t steps, double dt,
                             int8_t* out_s, double* out_prob, double* out_entropy, double* out_energy_trace,
                             double* out_energy, double* out_expected_energy);

/* tessera_sv_anneal that also returns the final state (n <= 30 here, 2^n
 * complex amplitudes interleaved re/im in out_state, bit i of the index = spin i
 * set to +1): the amplitudes upstream quantum.rs:78-87 holds after the last step.
 * Scalars as tessera_sv_anneal (nullable). */
int tessera_sv_anneal_state(const double* h, size_t n,
                            const size_t* ei, const size_t* ej, const double* ew, size_t m,
                            const double* schedule_a, const double* schedule_b, size_t steps, double dt,
                            int8_t* out_s, double* out_expected_energy, double* out_entropy, double* out_state);

/*
 * Drop-in for upstream rust/src/capi.rs:107-143 (the ctypes path of
 * python/test_ffi.py and python/tessera_cam.py): state-vector anneal on the
 * smooth cos^2/sin^2 schedule with steps = max(steps, 2). Returns the energy
 * of the readout; 0.0 when h or out_s is null; NaN when n > 24, as upstream.
 */
double tessera_solve_qa(const double* h, size_t n,
                        const size_t* ei, const size_t* ej, const double* ew, size_t m,
                        size_t steps, double dt, int8_t* out_s, double* out_prob,
                        double* out_entropy, double* out_energy);

/* ---- native TEBD session (no upstream counterpart: the runtime the entry
 * points above are built on, exposed for benchmarking and batched drivers) ---- */
typedef struct tessera_tebd tessera_tebd;

/* Allocate device state for n sites at bond dimension chi; model as above. */
int tessera_tebd_create(const double* h, size_t n, const size_t* ei, const size_t* ej,
                        const double* ew, size_t m, size_t chi, double dt, int zz_substeps,
                        tessera_tebd** out);
/* |+>^n product state (mps.cpp:9-17). */
int tessera_tebd_reset(tessera_tebd* t);
/* Warm start: the product of the single-spin ground states of
 * -(X_i + bias_i Z_i), the ground state at s = 0 of an anneal run with the
 * same bias (tessera_tebd_set_bias); bias = NULL gives |+>^n. */
int tessera_tebd_reset_warm(tessera_tebd* t, const double* bias);
/* Benchmark state: random site tensors with every bond at its capacity
 * min(chi, 2^b, 2^(n-b)), then one left-to-right sweep of identity two-site
 * updates that leaves every site but the last an isometry and gives bond b the
 * weights w_k ~ exp(-k/decay) (sum w_k^2 = 1): a saturated MPS whose Trotter
 * steps truncate little, like the saturated phase of a real anneal. */
int tessera_tebd_randomize(tessera_tebd* t, uint64_t seed, double decay);
/* Enqueue `count` Trotter steps with schedule values a[k], b[k] (no host sync);
 * qubit-0 entropy of each step is recorded on the device when record != 0. */
int tessera_tebd_run(tessera_tebd* t, const double* a, const double* b, size_t count, int record);
/* Wait for queued work; copy the recorded entropy trace (nullable), the max
 * discarded weight and the current bond dimensions (n+1 values, nullable). */
int tessera_tebd_sync(tessera_tebd* t, double* out_entropy, size_t cap, double* out_max_trunc,
                      int32_t* out_bonds);
/* Readout: spins, energy, expected energy, prob (as tessera_qa_anneal_ex). */
int tessera_tebd_readout(tessera_tebd* t, int8_t* out_s, double* out_scalars);
/* Per-spin expectations <Z_i> = <psi|Z_i|psi> / <psi|psi> of the current state
 * (n values; the quantities whose signs the readout takes, mps.cpp:369-372). */
int tessera_tebd_expect_z(tessera_tebd* t, double* out_z);
/* Qubit-0 entanglement entropy of the current state (mps.cpp:218-291). */
int tessera_tebd_entropy_now(tessera_tebd* t, double* out_entropy);
/* Per-spin longitudinal warm-start field for the steps enqueued from now on:
 * the single-site generator becomes -A X - (B h_i + A bias_i) Z (n values);
 * NULL removes it (the plain anneal). */
int tessera_tebd_set_bias(tessera_tebd* t, const double* bias);
/* Dense amplitudes (n <= 24): out[2*x], out[2*x+1] = re, im of psi(x), where bit
 * i of x is the physical index of site i (0 <-> spin +1). */
int tessera_tebd_dense(tessera_tebd* t, double* out);
/* Diagnostic: theta of sites (p, p+1) under the identity gate, the matrix the
 * next update at p factorises: out_rows x out_cols (2Dl x 2Dr) complex values,
 * row-major, interleaved re/im; `cap` = complex entries available in out (out
 * may be NULL to query the shape). */
int tessera_tebd_theta(tessera_tebd* t, size_t p, double* out, size_t cap, int32_t* out_rows, int32_t* out_cols);
/* Diagnostics since creation: SVDs run, Jacobi sweeps they took, out_rows[2] =
 * column-pair visits and rotations weighted by column length (pair-rows; the
 * one-sided Jacobi does 16 flops per visited pair-row and 18 more per rotated
 * one), out_ms[2] = HIP-event milliseconds of the timed SVD launches and their
 * count (see tessera_tebd_set_timing). Any output may be NULL. */
int tessera_tebd_svd_stats(tessera_tebd* t, int64_t* out_calls, int64_t* out_sweeps, double* out_rows,
                           double* out_ms);
/* Record HIP events around every SVD launch from now on (on != 0); resets the
 * accumulated time. */
int tessera_tebd_set_timing(tessera_tebd* t, int on);
/* Lockstep executor for tessera_tebd_run: 0 = off (each update is its own
 * sequence of launches, the default), g > 0 = one cooperative launch per run
 * call (per chunk of recorded steps) in which up to g groups of workgroups run
 * the steps' two-site updates in super-steps of mutually independent updates,
 * so that consecutive Trotter steps overlap; -1 = as many groups as fit on the
 * device. Single-site half layers are folded into the two-site gates; the
 * per-step entropy is taken from site snapshots. Results equal the launch-per-
 * update path up to rounding. */
int tessera_tebd_set_executor(tessera_tebd* t, int groups);
/* tessera_tebd_set_executor with the workgroups per group of the grouped layout:
 * 0 = one per block pair of a tournament round (the default), w > 0 = at most w,
 * each taking several pairs per round (fewer workgroups in every barrier). In a
 * batched run the first session's setting applies. */
int tessera_tebd_set_executor_shape(tessera_tebd* t, int groups, int workers);
/* Householder LQ preconditioning of the multi-workgroup Jacobi SVD (on != 0, the
 * default) or Jacobi directly on theta (0): the same factorisation, a different
 * sweep count; a knob of the autotuner. In a batched run the first session's
 * setting applies. */
int tessera_tebd_set_svd_lq(tessera_tebd* t, int on);
/* Batched anneal (no upstream counterpart; the batched scheduler): advance
 * count sessions together, session i by steps[i] Trotter steps with schedule
 * values a[i][k], b[i][k], in one cooperative launch whose super-steps take
 * independent two-site updates from all sessions (each session's gate order is
 * kept). groups as for tessera_tebd_set_executor (-1: as many as fit). No
 * entropy is recorded; read each session out with tessera_tebd_readout. */
int tessera_tebd_batch_run(tessera_tebd* const* sessions, size_t count, const double* const* a,
                           const double* const* b, const size_t* steps, int groups);
/* tessera_tebd_batch_run with the executor layout chosen explicitly: mode 0 =
 * automatic (solo for four or more sessions with chi <= 32, grouped otherwise), 1 =
 * grouped (each update's SVD spread over a group of workgroups that synchronise
 * on the group's own barrier: few, large updates), 2 = solo (one workgroup per update, one grid
 * barrier per super-step, up to one update per resident workgroup, thetas up
 * to 64 x 64 or 128 x 32 factorised in LDS: many independent small updates, as in batches
 * of small instances). groups caps the
 * concurrent updates (-1: as many as fit). Results agree to rounding. */
int tessera_tebd_batch_run_ex(tessera_tebd* const* sessions, size_t count, const double* const* a,
                              const double* const* b, const size_t* steps, int groups, int mode);
/* Per-instance wall time inside the last executor launch (single or batched)
 * that advanced this session: device-clock milliseconds from the launch start
 * to the done flag of the session's last update (dataflow order, the default)
 * or to the end of the super-step holding it (0 if that launch gave it no
 * update). No upstream counterpart: the batched scheduler's per-instance
 * latency. */
int tessera_tebd_last_run_ms(tessera_tebd* t, double* out_ms);
/* Lockstep executor totals since creation: super-steps run and two-site
 * updates placed in them (updates / (super-steps * groups) = slot fill). */
int tessera_tebd_exec_stats(tessera_tebd* t, int64_t* out_supersteps, int64_t* out_updates);
/* Diagnostics (no upstream counterpart): phase profile of the lockstep
 * executor's first workgroup since creation, first `count` of 96 slots: ms in
 * 0 snapshots + theta, 1 LQ preconditioning, 2 Jacobi sweeps, 3 whole SVD,
 * 4 wait after the SVD, 5 reconstruction, 6 wait at the super-step end (in the
 * dataflow order: claims from the update queue and waits for predecessors),
 * 8 waits at tournament-round barriers, 10 whole launches; counts in 7
 * updates run by its group, 9 tournament rounds, 11 Jacobi sweeps; ms of the
 * rounds' parts: 12 block loads, 13 in-block pairs, 14 cross pairs,
 * 15 write-back, 16 publishing the round (point-to-point rounds); counts per
 * Jacobi sweep index s < 24 (the last index also takes later sweeps):
 * 17 + s rotations, 41 + s rounds with a rotation above the stop ratio,
 * 65 + s SVDs that ran sweep s; shader cycles of thread 0 in the parts of the
 * row-layout rotation steps: 89 products and wave reduction, 90 barrier,
 * 91 totals and rotations, 92 rotation broadcast, 93 application (recorded
 * with TESSERA_EXEC_STEP_PROFILE=1 set when the session is created); ms in
 * 94 whole tournament rounds, 95 the flag checks between sweeps. */
int tessera_tebd_exec_profile(tessera_tebd* t, double* out, size_t count);
/* Number of two-site updates one Trotter step performs for this model. */
size_t tessera_tebd_updates_per_step(const tessera_tebd* t);
void tessera_tebd_destroy(tessera_tebd* t);

/* ---- learned guidance network (no upstream counterpart; the structure of
 * python/gnn_guide.py:81-103 IsingGNN is the template) ----
 * A 3-layer message-passing network of width 64 over Ising graphs, BF16 on
 * MFMA with FP32 accumulation and FP32 master weights (Adam). Node features
 * (h_i, deg_i / 3), edge feature J_ij. Outputs: per-spin z_i in (-1, 1), the
 * estimate of the final <Z_i> of the anneal (the warm-start bias), and per
 * graph s_g in (0, 1), the estimate of the schedule fraction at which the
 * qubit-0 entanglement entropy peaks (the schedule stretch is derived from it).
 * Loss: mean_i (z_i - tz_i)^2 + mean_g (s_g - ts_g)^2. */
typedef struct tessera_gnn tessera_gnn;

/* Fresh network: uniform(+-1/sqrt(fan_in)) weights from `seed`, zero biases. */
int tessera_gnn_create(uint64_t seed, tessera_gnn** out);
/* Number of FP32 parameters (the flat layout of get/set_params). */
size_t tessera_gnn_param_count(void);
int tessera_gnn_get_params(tessera_gnn* g, float* out);
/* Replace the master weights (and their BF16 copies); resets the Adam state. */
int tessera_gnn_set_params(tessera_gnn* g, const float* in);
/* Device-resident dataset of `graphs` graphs. Graph k owns nodes
 * node_off[k]..node_off[k+1]-1 (at least one); node_feat holds 2 floats per
 * node; node i's neighbours are adj_nbr[adj_off[i]..adj_off[i+1]-1] (global
 * node ids inside the same graph; list both directions of every edge) with
 * couplings adj_j. z_target (per node) and s_target (per graph) may both be
 * NULL for inference only. */
int tessera_gnn_set_data(tessera_gnn* g, size_t graphs, const int32_t* node_off, const float* node_feat,
                         const int32_t* adj_off, const int32_t* adj_nbr, const float* adj_j,
                         const float* z_target, const float* s_target);
/* Forward pass over graphs [g0, g1): out_z per node of the range, out_s per graph
 * (either may be NULL). */
int tessera_gnn_predict(tessera_gnn* g, size_t g0, size_t g1, float* out_z, float* out_s);
/* Loss of graphs [g0, g1) and its gradient (param_count floats, nullable); no update. */
int tessera_gnn_grad(tessera_gnn* g, size_t g0, size_t g1, float* out_grad, double* out_loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the loss of graphs [g0, g1);
 * out_loss (nullable) receives the loss before the step. */
int tessera_gnn_train_step(tessera_gnn* g, size_t g0, size_t g1, float lr, double* out_loss);
void tessera_gnn_destroy(tessera_gnn* g);

/* Text of the last error on this thread ("" if none). */
const char* tessera_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TESSERA_QA_H */
