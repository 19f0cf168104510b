// extern "C" entry points of libtessera_qa.so (declared in include/tessera_qa.h).
// Every entry point catches all exceptions: nothing unwinds across the C ABI.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <set>
#include <string>

#include "../../include/tessera_qa.h"
#include "common.hpp"
#include "gnn.hpp"
#include "sv.hpp"
#include "tebd.hpp"

namespace tq {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {

// tessera_qa_anneal(_ex): chains whose bond capacity reaches this run in the
// lockstep executor (at n = 128, chi = 256 about 7x the per-update path)
constexpr int kExecutorMinBond = 128;

bool finite_all(const double* p, std::size_t n) {
    for (std::size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// Shared validation of the model arguments; throws tq::Error.
void check_model(const double* h, std::size_t n, const std::size_t* ei, const std::size_t* ej, const double* ew,
                 std::size_t m, const double* a, const double* b, std::size_t steps) {
    if (!h || n == 0) throw Error(TESSERA_ERR_ARG, "h is null or n == 0");
    if (m > 0 && (!ei || !ej || !ew)) throw Error(TESSERA_ERR_ARG, "edge arrays are null with m > 0");
    if (steps > 0 && (!a || !b)) throw Error(TESSERA_ERR_ARG, "schedule arrays are null");
    for (std::size_t k = 0; k < m; ++k)
        if (ei[k] >= n || ej[k] >= n) throw Error(TESSERA_ERR_ARG, "edge index >= n");
    if (!finite_all(h, n) || (m && !finite_all(ew, m)) || (steps && (!finite_all(a, steps) || !finite_all(b, steps))))
        throw Error(TESSERA_ERR_NONFINITE, "non-finite value in h, J or schedule");
}

TebdModel make_model(const double* h, std::size_t n, const std::size_t* ei, const std::size_t* ej, const double* ew,
                     std::size_t m, const double* bias) {
    TebdModel md;
    md.n = (int)n;
    md.h.assign(h, h + n);
    md.ei.resize(m);
    md.ej.resize(m);
    for (std::size_t k = 0; k < m; ++k) {  // narrowed to int as ffi.cpp:19-21
        md.ei[k] = (int)ei[k];
        md.ej[k] = (int)ej[k];
    }
    md.ew.assign(ew, ew + m);
    if (bias) md.bias.assign(bias, bias + n);
    return md;
}

int status_of(const std::exception& e) {
    if (auto* te = dynamic_cast<const Error*>(&e)) return te->status;
    return TESSERA_ERR_DEVICE;
}

}  // namespace
}  // namespace tq

struct tessera_tebd {
    std::unique_ptr<tq::Tebd> eng;
    int recorded = 0;
};

struct tessera_gnn {
    std::unique_ptr<tq::Gnn> net;
};

using namespace tq;

extern "C" {

const char* tessera_last_error(void) { return tq::last_error(); }

int tessera_qa_anneal_ex(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew, size_t m,
                         const double* schedule_a, const double* schedule_b, size_t steps, double dt, size_t chi,
                         int zz_substeps, const double* bias, int8_t* out_s, double* out_scalars,
                         double* out_entropy, int32_t* out_max_bond) {
    try {
        set_last_error("");
        check_model(h, n, ei, ej, ew, m, schedule_a, schedule_b, steps);
        if (!out_s || !out_scalars) throw Error(TESSERA_ERR_ARG, "out_s or out_scalars is null");
        if (!std::isfinite(dt)) throw Error(TESSERA_ERR_NONFINITE, "dt is not finite");
        if (bias && !finite_all(bias, n)) throw Error(TESSERA_ERR_NONFINITE, "non-finite bias");
        if (chi == 0) throw Error(TESSERA_ERR_ARG, "chi == 0");
        if (steps > (size_t)INT32_MAX) throw Error(TESSERA_ERR_ARG, "too many steps");
        Tebd t(make_model(h, n, ei, ej, ew, m, bias), (int)chi, dt, zz_substeps);
        // bonds that can reach 128 or more (thetas of 256 x 256 and up): the lockstep
        // executor, every launch of up to 64 steps one cooperative kernel; smaller
        // chains: the per-update path (one captured Trotter step replayed per step,
        // where every update's SVD takes a single-workgroup kernel)
        if (t.max_capacity() >= kExecutorMinBond) t.set_executor(-1, 0);
        t.run(schedule_a, schedule_b, (int)steps, 0, out_max_bond ? 0 : -1);
        t.sync();
        if (out_max_bond && steps) {
            const std::vector<int> mb = t.max_bonds((int)steps);
            std::memcpy(out_max_bond, mb.data(), steps * sizeof(int32_t));
        }
        if (out_entropy && steps) {
            const std::vector<double> ent = t.entropy((int)steps);
            std::memcpy(out_entropy, ent.data(), steps * sizeof(double));
        }
        const Tebd::Readout r = t.readout();
        std::memcpy(out_s, r.spins.data(), n);
        out_scalars[0] = r.energy;
        out_scalars[1] = r.expected_energy;
        out_scalars[2] = r.prob;
        out_scalars[3] = t.max_trunc();
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    } catch (...) {
        set_last_error("unknown exception");
        return TESSERA_ERR_DEVICE;
    }
}

double tessera_qa_anneal(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew, size_t m,
                         const double* schedule_a, const double* schedule_b, size_t steps, size_t chi, int8_t* out_s,
                         double* out_entropy) {
    double sc[4] = {0, 0, 0, 0};
    const int st = tessera_qa_anneal_ex(h, n, ei, ej, ew, m, schedule_a, schedule_b, steps, 0.1, chi, 1, nullptr,
                                        out_s, sc, out_entropy, nullptr);
    if (st != TESSERA_OK) {
        if (out_s)
            for (size_t i = 0; i < n; ++i) out_s[i] = 1;
        return std::nan("");
    }
    return sc[0];
}

}  // extern "C"

namespace {
// tessera_sv_anneal and tessera_sv_anneal_static: the same anneal, device
// buffers allocated per call or carved from the static scratch.
int sv_entry(bool static_scratch, const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew,
             size_t m, const double* schedule_a, const double* schedule_b, size_t steps, double dt, int8_t* out_s,
             double* out_prob, double* out_entropy, double* out_energy_trace, double* out_energy,
             double* out_expected_energy) {
    try {
        set_last_error("");
        check_model(h, n, ei, ej, ew, m, schedule_a, schedule_b, steps);
        if (!out_s) throw Error(TESSERA_ERR_ARG, "out_s is null");
        if (steps < 1) throw Error(TESSERA_ERR_ARG, "steps < 1");
        if (steps > (size_t)INT32_MAX || m > (size_t)INT32_MAX) throw Error(TESSERA_ERR_ARG, "too many steps or edges");
        if (!std::isfinite(dt)) throw Error(TESSERA_ERR_NONFINITE, "dt is not finite");
        if (n > 36) throw Error(TESSERA_ERR_TOO_LARGE, "n > 36 does not fit one device");
        SvProblem pb;
        pb.n = (int)n;
        pb.h.assign(h, h + n);
        pb.ei.resize(m);
        pb.ej.resize(m);
        for (size_t k = 0; k < m; ++k) {
            pb.ei[k] = (int)ei[k];
            pb.ej[k] = (int)ej[k];
        }
        pb.ew.assign(ew, ew + m);
        pb.a.assign(schedule_a, schedule_a + steps);
        pb.b.assign(schedule_b, schedule_b + steps);
        pb.steps = (int)steps;
        pb.dt = dt;
        const SvOutcome r = static_scratch ? sv_anneal_static(pb) : sv_anneal_device(pb, false);
        std::memcpy(out_s, r.spins.data(), n);
        if (out_prob) *out_prob = r.success_prob;
        if (out_entropy) std::memcpy(out_entropy, r.entropy.data(), steps * sizeof(double));
        if (out_energy_trace) std::memcpy(out_energy_trace, r.energy_trace.data(), steps * sizeof(double));
        if (out_energy) *out_energy = r.energy;
        if (out_expected_energy) *out_expected_energy = r.expected_energy;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    } catch (...) {
        set_last_error("unknown exception");
        return TESSERA_ERR_DEVICE;
    }
}
}  // namespace

extern "C" {

int tessera_sv_anneal(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew, size_t m,
                      const double* schedule_a, const double* schedule_b, size_t steps, double dt, int8_t* out_s,
                      double* out_prob, double* out_entropy, double* out_energy_trace, double* out_energy,
                      double* out_expected_energy) {
    return sv_entry(false, h, n, ei, ej, ew, m, schedule_a, schedule_b, steps, dt, out_s, out_prob, out_entropy,
                    out_energy_trace, out_energy, out_expected_energy);
}

int tessera_sv_anneal_static(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew,
                             size_t m, const double* schedule_a, const double* schedule_b, size_t steps, double dt,
                             int8_t* out_s, double* out_prob, double* out_entropy, double* out_energy_trace,
                             double* out_energy, double* out_expected_energy) {
    return sv_entry(true, h, n, ei, ej, ew, m, schedule_a, schedule_b, steps, dt, out_s, out_prob, out_entropy,
                    out_energy_trace, out_energy, out_expected_energy);
}

int tessera_sv_anneal_state(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew,
                            size_t m, const double* schedule_a, const double* schedule_b, size_t steps, double dt,
                            int8_t* out_s, double* out_expected_energy, double* out_entropy, double* out_state) {
    try {
        set_last_error("");
        check_model(h, n, ei, ej, ew, m, schedule_a, schedule_b, steps);
        if (!out_s || !out_state) throw Error(TESSERA_ERR_ARG, "out_s or out_state is null");
        if (steps < 1) throw Error(TESSERA_ERR_ARG, "steps < 1");
        if (!std::isfinite(dt)) throw Error(TESSERA_ERR_NONFINITE, "dt is not finite");
        if (n > 30) throw Error(TESSERA_ERR_TOO_LARGE, "state download limited to n <= 30");
        SvProblem pb;
        pb.n = (int)n;
        pb.h.assign(h, h + n);
        pb.ei.resize(m);
        pb.ej.resize(m);
        for (size_t k = 0; k < m; ++k) {
            pb.ei[k] = (int)ei[k];
            pb.ej[k] = (int)ej[k];
        }
        pb.ew.assign(ew, ew + m);
        pb.a.assign(schedule_a, schedule_a + steps);
        pb.b.assign(schedule_b, schedule_b + steps);
        pb.steps = (int)steps;
        pb.dt = dt;
        const SvOutcome r = sv_anneal_device(pb, true);
        std::memcpy(out_s, r.spins.data(), n);
        if (out_expected_energy) *out_expected_energy = r.expected_energy;
        if (out_entropy) std::memcpy(out_entropy, r.entropy.data(), steps * sizeof(double));
        std::memcpy(out_state, r.state.data(), r.state.size() * sizeof(double));
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    } catch (...) {
        set_last_error("unknown exception");
        return TESSERA_ERR_DEVICE;
    }
}

// rust/src/capi.rs:107-143 semantics on the device state vector.
double tessera_solve_qa(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew, size_t m,
                        size_t steps, double dt, int8_t* out_s, double* out_prob, double* out_entropy,
                        double* out_energy) {
    if (!h || !out_s) return 0.0;
    if (n > 24) return std::nan("");
    const size_t st = steps < 2 ? 2 : steps;
    std::vector<double> a(st), b(st);
    for (size_t k = 0; k < st; ++k) {  // schedule.rs:35-46
        const double s = (double)k / (double)(st - 1);
        const double th = 1.5707963267948966 * s;
        const double c = std::cos(th), sn = std::sin(th);
        a[k] = c * c;
        b[k] = sn * sn;
    }
    // capi.rs:32-39 treats a null ei as "no couplings"
    const size_t mm = (m > 0 && ei) ? m : 0;
    double energy = std::nan("");
    const int rc = tessera_sv_anneal(h, n, mm ? ei : nullptr, mm ? ej : nullptr, mm ? ew : nullptr, mm, a.data(),
                                     b.data(), st, dt, out_s, out_prob, out_entropy, out_energy, &energy, nullptr);
    if (rc != TESSERA_OK) return std::nan("");
    return energy;
}

int tessera_tebd_create(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew, size_t m,
                        size_t chi, double dt, int zz_substeps, tessera_tebd** out) {
    try {
        set_last_error("");
        if (!out) throw Error(TESSERA_ERR_ARG, "out is null");
        check_model(h, n, ei, ej, ew, m, nullptr, nullptr, 0);
        if (chi == 0) throw Error(TESSERA_ERR_ARG, "chi == 0");
        auto* t = new tessera_tebd;
        try {
            t->eng = std::make_unique<Tebd>(make_model(h, n, ei, ej, ew, m, nullptr), (int)chi, dt, zz_substeps);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_reset(tessera_tebd* t) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        t->eng->reset_product();
        t->recorded = 0;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_randomize(tessera_tebd* t, uint64_t seed, double decay) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        if (!(decay > 0)) throw Error(TESSERA_ERR_ARG, "decay must be > 0");
        t->eng->randomize(seed, decay);
        t->recorded = 0;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_run(tessera_tebd* t, const double* a, const double* b, size_t count, int record) {
    try {
        if (!t || (count && (!a || !b))) throw Error(TESSERA_ERR_ARG, "null argument");
        t->eng->run(a, b, (int)count, record ? t->recorded : -1);
        if (record) t->recorded += (int)count;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_sync(tessera_tebd* t, double* out_entropy, size_t cap, double* out_max_trunc, int32_t* out_bonds) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        t->eng->sync();
        if (out_entropy && cap) {
            const int cnt = (int)std::min<size_t>(cap, (size_t)t->recorded);
            const std::vector<double> e = t->eng->entropy(cnt);
            std::memcpy(out_entropy, e.data(), cnt * sizeof(double));
        }
        if (out_max_trunc) *out_max_trunc = t->eng->max_trunc();
        if (out_bonds) {
            const std::vector<int> d = t->eng->bonds();
            for (size_t i = 0; i < d.size(); ++i) out_bonds[i] = d[i];
        }
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_readout(tessera_tebd* t, int8_t* out_s, double* out_scalars) {
    try {
        if (!t || !out_s || !out_scalars) throw Error(TESSERA_ERR_ARG, "null argument");
        const Tebd::Readout r = t->eng->readout();
        std::memcpy(out_s, r.spins.data(), r.spins.size());
        out_scalars[0] = r.energy;
        out_scalars[1] = r.expected_energy;
        out_scalars[2] = r.prob;
        out_scalars[3] = t->eng->max_trunc();
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_reset_warm(tessera_tebd* t, const double* bias) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        if (bias && !finite_all(bias, (std::size_t)t->eng->n())) throw Error(TESSERA_ERR_NONFINITE, "non-finite bias");
        t->eng->reset_product(bias);
        t->recorded = 0;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_expect_z(tessera_tebd* t, double* out_z) {
    try {
        if (!t || !out_z) throw Error(TESSERA_ERR_ARG, "null argument");
        const Tebd::Readout r = t->eng->readout();
        std::memcpy(out_z, r.z.data(), r.z.size() * sizeof(double));
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_entropy_now(tessera_tebd* t, double* out_entropy) {
    try {
        if (!t || !out_entropy) throw Error(TESSERA_ERR_ARG, "null argument");
        *out_entropy = t->eng->entropy_now();
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_set_bias(tessera_tebd* t, const double* bias) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        if (bias && !finite_all(bias, (std::size_t)t->eng->n())) throw Error(TESSERA_ERR_NONFINITE, "non-finite bias");
        t->eng->set_bias(bias);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_dense(tessera_tebd* t, double* out) {
    try {
        if (!t || !out) throw Error(TESSERA_ERR_ARG, "null argument");
        const std::vector<double> d = t->eng->dense();
        std::memcpy(out, d.data(), d.size() * sizeof(double));
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_theta(tessera_tebd* t, size_t p, double* out, size_t cap, int32_t* out_rows, int32_t* out_cols) {
    try {
        if (!t || !out_rows || !out_cols) throw Error(TESSERA_ERR_ARG, "null argument");
        int rows = 0, cols = 0;
        const std::vector<double> th = t->eng->theta((int)p, rows, cols);
        *out_rows = rows;
        *out_cols = cols;
        if (out) {
            if (cap < th.size() / 2) throw Error(TESSERA_ERR_ARG, "out too small for theta");
            std::memcpy(out, th.data(), th.size() * sizeof(double));
        }
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_svd_stats(tessera_tebd* t, int64_t* out_calls, int64_t* out_sweeps, double* out_rows,
                           double* out_ms) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        const SvdStats st = t->eng->svd_stats();
        if (out_calls) *out_calls = st.calls;
        if (out_sweeps) *out_sweeps = st.sweeps;
        if (out_rows) {
            out_rows[0] = st.visited_rows;
            out_rows[1] = st.rotated_rows;
        }
        if (out_ms) {
            out_ms[0] = st.svd_ms;
            out_ms[1] = (double)st.timed_launches;
        }
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_set_timing(tessera_tebd* t, int on) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        t->eng->set_timing(on != 0);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_set_executor(tessera_tebd* t, int groups) {
    return tessera_tebd_set_executor_shape(t, groups, 0);
}

int tessera_tebd_set_svd_lq(tessera_tebd* t, int on) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        t->eng->set_svd_lq(on != 0);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_set_executor_shape(tessera_tebd* t, int groups, int workers) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        if (groups < -1) throw Error(TESSERA_ERR_ARG, "groups must be >= -1");
        if (workers < 0) throw Error(TESSERA_ERR_ARG, "workers must be >= 0");
        t->eng->set_executor(groups, workers);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_batch_run(tessera_tebd* const* sessions, size_t count, const double* const* a,
                           const double* const* b, const size_t* steps, int groups) {
    return tessera_tebd_batch_run_ex(sessions, count, a, b, steps, groups, 0);
}

int tessera_tebd_batch_run_ex(tessera_tebd* const* sessions, size_t count, const double* const* a,
                              const double* const* b, const size_t* steps, int groups, int mode) {
    try {
        if (mode < 0 || mode > 2) throw Error(TESSERA_ERR_ARG, "mode must be 0, 1 or 2");
        if (!sessions || (count && (!a || !b || !steps))) throw Error(TESSERA_ERR_ARG, "null argument");
        if (groups < -1) throw Error(TESSERA_ERR_ARG, "groups must be >= -1");
        // validate every entry before any session advances: an error leaves all
        // sessions as they were
        std::set<const Tebd*> seen;
        for (size_t i = 0; i < count; ++i) {
            if (!sessions[i]) throw Error(TESSERA_ERR_ARG, "null session");
            if (steps[i] > (size_t)INT32_MAX) throw Error(TESSERA_ERR_ARG, "too many steps");
            if (steps[i] && (!a[i] || !b[i])) throw Error(TESSERA_ERR_ARG, "null schedule");
            for (size_t k = 0; k < steps[i]; ++k)
                if (!std::isfinite(a[i][k]) || !std::isfinite(b[i][k])) throw Error(TESSERA_ERR_NONFINITE, "non-finite schedule");
            if (!seen.insert(sessions[i]->eng.get()).second) throw Error(TESSERA_ERR_ARG, "session listed twice");
        }
        std::vector<Tebd*> ts;
        std::vector<const double*> as, bs;
        std::vector<int> cs;
        for (size_t i = 0; i < count; ++i) {
            if (sessions[i]->eng->n() < 2) {  // nothing two-site to batch: plain path
                sessions[i]->eng->run(a[i], b[i], (int)steps[i], -1);
                continue;
            }
            ts.push_back(sessions[i]->eng.get());
            as.push_back(a[i]);
            bs.push_back(b[i]);
            cs.push_back((int)steps[i]);
        }
        Tebd::run_lockstep(ts, as, bs, cs, -1, groups, mode);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_exec_profile(tessera_tebd* t, double* out, size_t count) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        if (!out && count) throw Error(TESSERA_ERR_ARG, "null output");
        const std::vector<double> prof = t->eng->exec_profile();
        for (size_t q = 0; q < count && q < prof.size(); ++q) out[q] = prof[q];
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_last_run_ms(tessera_tebd* t, double* out_ms) {
    try {
        if (!t || !out_ms) throw Error(TESSERA_ERR_ARG, "null argument");
        *out_ms = t->eng->last_run_ms();
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_tebd_exec_stats(tessera_tebd* t, int64_t* out_supersteps, int64_t* out_updates) {
    try {
        if (!t) throw Error(TESSERA_ERR_ARG, "null session");
        const SvdStats st = t->eng->svd_stats();
        if (out_supersteps) *out_supersteps = st.supersteps;
        if (out_updates) *out_updates = st.exec_updates;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_create(uint64_t seed, tessera_gnn** out) {
    try {
        if (!out) throw Error(TESSERA_ERR_ARG, "null argument");
        *out = nullptr;
        auto g = std::make_unique<tessera_gnn>();
        g->net = std::make_unique<Gnn>(seed);
        *out = g.release();
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

size_t tessera_gnn_param_count(void) { return (size_t)GnnLayout::COUNT; }

int tessera_gnn_get_params(tessera_gnn* g, float* out) {
    try {
        if (!g || !out) throw Error(TESSERA_ERR_ARG, "null argument");
        g->net->get_params(out);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_set_params(tessera_gnn* g, const float* in) {
    try {
        if (!g || !in) throw Error(TESSERA_ERR_ARG, "null argument");
        g->net->set_params(in);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_set_data(tessera_gnn* g, size_t graphs, const int32_t* node_off, const float* node_feat,
                         const int32_t* adj_off, const int32_t* adj_nbr, const float* adj_j,
                         const float* z_target, const float* s_target) {
    try {
        if (!g || !node_off || !node_feat || !adj_off) throw Error(TESSERA_ERR_ARG, "null argument");
        if (graphs == 0 || graphs > (size_t)INT32_MAX) throw Error(TESSERA_ERR_ARG, "graphs out of range");
        if ((z_target == nullptr) != (s_target == nullptr))
            throw Error(TESSERA_ERR_ARG, "z_target and s_target must both be set or both be NULL");
        if (node_off[graphs] <= 0) throw Error(TESSERA_ERR_ARG, "no nodes");
        if (adj_off[node_off[graphs]] > 0 && (!adj_nbr || !adj_j)) throw Error(TESSERA_ERR_ARG, "null adjacency");
        g->net->set_data((int)graphs, node_off, node_feat, adj_off, adj_nbr, adj_j, z_target, s_target);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_predict(tessera_gnn* g, size_t g0, size_t g1, float* out_z, float* out_s) {
    try {
        if (!g) throw Error(TESSERA_ERR_ARG, "null network");
        if (g0 >= g1 || g1 > (size_t)g->net->graphs()) throw Error(TESSERA_ERR_ARG, "bad graph range");
        g->net->predict((int)g0, (int)g1, out_z, out_s);
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_grad(tessera_gnn* g, size_t g0, size_t g1, float* out_grad, double* out_loss) {
    try {
        if (!g) throw Error(TESSERA_ERR_ARG, "null network");
        if (g0 >= g1 || g1 > (size_t)g->net->graphs()) throw Error(TESSERA_ERR_ARG, "bad graph range");
        const double loss = g->net->grad((int)g0, (int)g1, out_grad);
        if (out_loss) *out_loss = loss;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

int tessera_gnn_train_step(tessera_gnn* g, size_t g0, size_t g1, float lr, double* out_loss) {
    try {
        if (!g) throw Error(TESSERA_ERR_ARG, "null network");
        if (g0 >= g1 || g1 > (size_t)g->net->graphs()) throw Error(TESSERA_ERR_ARG, "bad graph range");
        const double loss = g->net->train_step((int)g0, (int)g1, lr);
        if (out_loss) *out_loss = loss;
        return TESSERA_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return status_of(e);
    }
}

void tessera_gnn_destroy(tessera_gnn* g) { delete g; }

size_t tessera_tebd_updates_per_step(const tessera_tebd* t) { return t ? t->eng->updates_per_step() : 0; }

void tessera_tebd_destroy(tessera_tebd* t) { delete t; }

}  // extern "C"
