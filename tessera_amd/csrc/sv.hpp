// Exact state-vector engine (upstream rust/src/quantum.rs).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace tq {

struct SvProblem {
    int n = 0;
    std::vector<double> h;
    std::vector<int> ei, ej;
    std::vector<double> ew;
    std::vector<double> a, b;  // schedule, steps entries each
    int steps = 0;
    double dt = 0.0;
};

struct SvOutcome {  // quantum.rs:27-43
    std::vector<int8_t> spins;
    double energy = 0, expected_energy = 0, success_prob = 0;
    std::vector<double> entropy, energy_trace;
    long long best_index = 0;
    std::vector<double> state;  // interleaved re/im, only when requested
};

// Qubit-0 entropy from the reduced density matrix sums (quantum.rs:156-174).
__host__ __device__ double sv_entropy_from(double r00, double r11, double r01r, double r01i);

// Device memory of one anneal (state, traces, tables, per-block partials) comes
// from `arena`: nullptr allocates it for the call; a region-backed arena carves
// it from caller-provided memory and allocates nothing.
SvOutcome sv_anneal_device(const SvProblem& pb, bool keep_state, DevArena* arena = nullptr);

// tessera_sv_anneal without device allocation: every buffer of the anneal lives
// in one static device array of kSvStaticBytes per GPU (n <= kSvStaticMaxN).
constexpr int kSvStaticMaxN = 16;
constexpr int kSvStaticMaxSteps = 16384;
constexpr int kSvStaticMaxEdges = 4096;
constexpr std::size_t kSvStaticBytes = (std::size_t)4 << 20;
SvOutcome sv_anneal_static(const SvProblem& pb);

}  // namespace tq
