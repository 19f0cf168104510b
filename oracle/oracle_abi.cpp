// extern "C" surface of the oracle for ctypes (tests, smoke check, bench CPU
// baseline). TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <chrono>
#include <cstring>
#include <random>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

// Plain tessera_qa_anneal semantics (cuda/src/ffi.cpp:10-37): dt = 0.1, one ZZ slice.
double oracle_qa_anneal(const double* h, std::size_t n, const std::size_t* ei, const std::size_t* ej,
                        const double* ew, std::size_t m, const double* sa, const double* sb,
                        std::size_t steps, std::size_t chi, int8_t* out_s, double* out_entropy) {
    std::vector<double> hv(h, h + n), ewv(ew, ew + m), av(sa, sa + steps), bv(sb, sb + steps);
    std::vector<int> eiv(m), ejv(m);
    for (std::size_t k = 0; k < m; ++k) {
        eiv[k] = (int)ei[k];
        ejv[k] = (int)ej[k];
    }
    TebdResult r = tebd_anneal(hv, eiv, ejv, ewv, av, bv, 0.1, (int)chi, 1);
    for (std::size_t i = 0; i < n; ++i) out_s[i] = r.spins[i];
    if (out_entropy)
        for (std::size_t k = 0; k < steps; ++k) out_entropy[k] = k < r.entropy.size() ? r.entropy[k] : 0.0;
    return r.energy;
}

// Full TEBD run. out_scalars = {energy, expected_energy, prob, max_trunc}.
// Optional (nullable): per-step entropy, max bond, norm2, and the dense final
// state (2 * 2^n doubles, interleaved re/im; n <= 24). svd_mode 0 is the
// upstream Gram SVD (bit-exact), 1 the accurate one-sided Jacobi SVD.
// out_ambiguous[k] (nullable): updates of step k whose kept rank differs
// between cuts of 1e-13 and 1e-15, i.e. singular values at the rounding level
// of theta next to the 1e-14 cut.
int oracle_tebd(const double* h, std::size_t n, const std::size_t* ei, const std::size_t* ej,
                const double* ew, std::size_t m, const double* sa, const double* sb, std::size_t steps,
                double dt, std::size_t chi, int zz_substeps, int8_t* out_s, double* out_scalars,
                double* out_entropy, int* out_max_bond, double* out_norm2, double* out_state, int svd_mode,
                int* out_ambiguous) {
    if (!h || !out_s || !out_scalars || n == 0) return -1;
    if (out_state && n > 24) return -2;
    std::vector<double> hv(h, h + n), ewv(ew, ew + m), av(sa, sa + steps), bv(sb, sb + steps);
    std::vector<int> eiv(m), ejv(m);
    for (std::size_t k = 0; k < m; ++k) {
        eiv[k] = (int)ei[k];
        ejv[k] = (int)ej[k];
    }
    TebdResult r = tebd_anneal(hv, eiv, ejv, ewv, av, bv, dt, (int)chi, zz_substeps, out_norm2 != nullptr,
                               out_state != nullptr, svd_mode);
    for (std::size_t i = 0; i < n; ++i) out_s[i] = r.spins[i];
    out_scalars[0] = r.energy;
    out_scalars[1] = r.expected_energy;
    out_scalars[2] = r.prob;
    out_scalars[3] = r.max_trunc;
    for (std::size_t k = 0; k < steps; ++k) {
        if (out_entropy) out_entropy[k] = r.entropy[k];
        if (out_max_bond) out_max_bond[k] = r.max_bond[k];
        if (out_norm2) out_norm2[k] = r.norm2[k];
        if (out_ambiguous) out_ambiguous[k] = r.ambiguous[k];
    }
    if (out_state)
        for (std::size_t x = 0; x < r.state.size(); ++x) {
            out_state[2 * x] = r.state[x].real();
            out_state[2 * x + 1] = r.state[x].imag();
        }
    return 0;
}

// Exact state vector. out_scalars = {energy, expected_energy, success_prob}.
int oracle_sv(const double* h, std::size_t n, const std::size_t* ei, const std::size_t* ej,
              const double* ew, std::size_t m, const double* sa, const double* sb, std::size_t steps,
              double dt, int8_t* out_s, double* out_scalars, double* out_entropy, double* out_energy,
              double* out_state) {
    if (!h || !out_s || !out_scalars || n == 0 || n > 30) return -1;
    std::vector<double> hv(h, h + n), ewv(ew, ew + m), av(sa, sa + steps), bv(sb, sb + steps);
    std::vector<std::size_t> eiv(ei, ei + m), ejv(ej, ej + m);
    SvResult r = sv_anneal(hv, eiv, ejv, ewv, av, bv, dt, out_state != nullptr);
    for (std::size_t i = 0; i < n; ++i) out_s[i] = r.spins[i];
    out_scalars[0] = r.energy;
    out_scalars[1] = r.expected_energy;
    out_scalars[2] = r.success_prob;
    for (std::size_t k = 0; k < steps; ++k) {
        if (out_entropy) out_entropy[k] = r.entropy[k];
        if (out_energy) out_energy[k] = r.energy_trace[k];
    }
    if (out_state)
        for (std::size_t x = 0; x < r.re.size(); ++x) {
            out_state[2 * x] = r.re[x];
            out_state[2 * x + 1] = r.im[x];
        }
    return 0;
}

// Dense numerics for known-answer tests. Matrices are row-major interleaved re/im.
int oracle_herm_eig(const double* hm, std::size_t n, double* out_w, double* out_v) {
    Mat a(n, n);
    for (std::size_t i = 0; i < n * n; ++i) a.v[i] = cplx(hm[2 * i], hm[2 * i + 1]);
    std::vector<double> w;
    Mat v;
    herm_eig(a, w, v);
    for (std::size_t i = 0; i < n; ++i) out_w[i] = w[i];
    for (std::size_t i = 0; i < n * n; ++i) {
        out_v[2 * i] = v.v[i].real();
        out_v[2 * i + 1] = v.v[i].imag();
    }
    return 0;
}

int oracle_expm(const double* hm, std::size_t n, double coeff_re, double coeff_im, double* out) {
    Mat a(n, n);
    for (std::size_t i = 0; i < n * n; ++i) a.v[i] = cplx(hm[2 * i], hm[2 * i + 1]);
    Mat e = herm_expm(a, cplx(coeff_re, coeff_im));
    for (std::size_t i = 0; i < n * n; ++i) {
        out[2 * i] = e.v[i].real();
        out[2 * i + 1] = e.v[i].imag();
    }
    return 0;
}

// Thin Gram SVD of an r x c matrix: U (r x k), S (k), Vh (k x c), k = min(r, c).
int oracle_svd(const double* am, std::size_t r, std::size_t c, double* out_u, double* out_s, double* out_vh) {
    Mat a(r, c);
    for (std::size_t i = 0; i < r * c; ++i) a.v[i] = cplx(am[2 * i], am[2 * i + 1]);
    Mat u, vh;
    std::vector<double> s;
    gram_svd(a, u, s, vh);
    for (std::size_t i = 0; i < s.size(); ++i) out_s[i] = s[i];
    for (std::size_t i = 0; i < u.v.size(); ++i) {
        out_u[2 * i] = u.v[i].real();
        out_u[2 * i + 1] = u.v[i].imag();
    }
    for (std::size_t i = 0; i < vh.v.size(); ++i) {
        out_vh[2 * i] = vh.v[i].real();
        out_vh[2 * i + 1] = vh.v[i].imag();
    }
    return 0;
}

// CPU-baseline probe for one saturated two-site update at bond dimension chi:
// theta is a full-rank 2chi x 2chi complex matrix Q1 diag(exp(-k/40)) Q2^dagger
// (Q1, Q2 from Gram-Schmidt of seeded Gaussian matrices), factorised as the
// upstream update does (linalg.cpp:107-172): Gram matrix, 4chi x 4chi real
// embedding, cyclic Jacobi. The Jacobi runs at most `max_sweeps` sweeps.
// Returns the seconds of Gram + embedding; writes the sweeps run and the seconds
// of the Jacobi sweeps.
double oracle_update_probe(int chi, int max_sweeps, int* out_sweeps, double* out_sweep_seconds) {
    std::mt19937_64 rng(7);
    std::normal_distribution<double> g(0.0, 1.0);
    const std::size_t n = (std::size_t)chi * 2;
    auto unitary = [&]() {
        Mat q(n, n);
        for (auto& x : q.v) x = cplx(g(rng), g(rng));
        for (std::size_t j = 0; j < n; ++j) {
            for (std::size_t p = 0; p < j; ++p) {
                cplx d(0, 0);
                for (std::size_t i = 0; i < n; ++i) d += std::conj(q.at(i, p)) * q.at(i, j);
                for (std::size_t i = 0; i < n; ++i) q.at(i, j) -= d * q.at(i, p);
            }
            double nr = 0;
            for (std::size_t i = 0; i < n; ++i) nr += std::norm(q.at(i, j));
            nr = std::sqrt(nr);
            for (std::size_t i = 0; i < n; ++i) q.at(i, j) /= nr;
        }
        return q;
    };
    const Mat q1 = unitary(), q2 = unitary();
    Mat th(n, n);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            cplx acc(0, 0);
            for (std::size_t k = 0; k < n; ++k) acc += q1.at(i, k) * std::exp(-(double)k / 40.0) * std::conj(q2.at(j, k));
            th.at(i, j) = acc;
        }
    auto t0 = std::chrono::steady_clock::now();
    Mat gm(n, n);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            cplx acc(0, 0);
            for (std::size_t t = 0; t < n; ++t) acc += std::conj(th.at(t, i)) * th.at(t, j);
            gm.at(i, j) = acc;
        }
    const std::size_t mm = 2 * n;
    std::vector<double> emb(mm * mm, 0.0);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            emb[i * mm + j] = gm.at(i, j).real();
            emb[(i + n) * mm + (j + n)] = gm.at(i, j).real();
            emb[i * mm + (j + n)] = -gm.at(i, j).imag();
            emb[(i + n) * mm + j] = gm.at(i, j).imag();
        }
    auto t1 = std::chrono::steady_clock::now();
    std::vector<double> w, v;
    const int sw = sym_jacobi(emb, mm, w, v, max_sweeps);
    auto t2 = std::chrono::steady_clock::now();
    if (out_sweeps) *out_sweeps = sw;
    if (out_sweep_seconds) *out_sweep_seconds = std::chrono::duration<double>(t2 - t1).count();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"

extern "C" {
// Instance generator of the upstream MPS tests (cuda/tests/test_mps.cpp:100-111,
// 159-167): std::mt19937_64(seed), U(-1,1) for every h_i, then for each pair
// i < j one draw deciding the edge (> 0) and, if kept, one draw for J_ij.
// Consecutive calls with the same generator object are reproduced by `skip`
// (number of earlier instances drawn from the same stream). Returns m.
std::size_t oracle_upstream_instance(unsigned long long seed, int skip, int n, double* h,
                                     std::size_t* ei, std::size_t* ej, double* ew, std::size_t cap) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::size_t m = 0;
    for (int inst = 0; inst <= skip; ++inst) {
        m = 0;
        for (int i = 0; i < n; ++i) h[i] = u(rng);
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
                if (u(rng) > 0.0) {
                    const double w = u(rng);
                    if (m < cap) {
                        ei[m] = (std::size_t)i;
                        ej[m] = (std::size_t)j;
                        ew[m] = w;
                    }
                    ++m;
                }
    }
    return m;
}
}
