// Driver for the upstream TESSERA C++ engine (compiled from its own sources by
// oracle/Makefile into oracle/_ref/). It only calls the engine's public API
// (cuda/include/mps.hpp, cuda/include/linalg.hpp) and prints results with full
// precision, so that tests/golden/make_fixtures.py can record them as vectors.
// Never shipped, never linked into the product.
//
// Modes (instance text on stdin, format documented in tests/golden/README.md):
//   anneal          mps_anneal(...) outputs                 (mps.cpp:295-381)
//   trace           same driver loop through the public MPS methods, printing
//                   max_bond() and entropy after every step  (mps.cpp:309-356)
//   svd N REPS      time tessera::svd on an N x N complex matrix with singular
//                   values exp(-k/40)                        (linalg.cpp:107-172)
//   linalg          known-answer outputs of hermitian_eig / hermitian_expm / svd
#include "mps.hpp"
#include "linalg.hpp"
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace tessera;

struct Inst {
    int n = 0, m = 0, steps = 0, chi = 0, ns = 1;
    double dt = 0.1;
    std::vector<double> h, ew, a, b;
    std::vector<int> ei, ej;
};

static bool read_inst(Inst& I) {
    if (std::scanf("%d %d %d %d %d %lf", &I.n, &I.m, &I.steps, &I.chi, &I.ns, &I.dt) != 6) return false;
    I.h.resize(I.n); I.ei.resize(I.m); I.ej.resize(I.m); I.ew.resize(I.m);
    I.a.resize(I.steps); I.b.resize(I.steps);
    for (auto& x : I.h) if (std::scanf("%lf", &x) != 1) return false;
    for (int e = 0; e < I.m; ++e)
        if (std::scanf("%d %d %lf", &I.ei[e], &I.ej[e], &I.ew[e]) != 3) return false;
    for (auto& x : I.a) if (std::scanf("%lf", &x) != 1) return false;
    for (auto& x : I.b) if (std::scanf("%lf", &x) != 1) return false;
    return true;
}

static void print_vec(const char* key, const std::vector<double>& v) {
    std::printf("%s", key);
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}

static int mode_anneal() {
    Inst I;
    if (!read_inst(I)) { std::fprintf(stderr, "bad instance\n"); return 2; }
    auto t0 = std::chrono::steady_clock::now();
    AnnealResult r = mps_anneal(I.h, I.ei, I.ej, I.ew, I.a, I.b, I.dt, I.chi, I.ns);
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("spins");
    for (int8_t s : r.spins) std::printf(" %d", (int)s);
    std::printf("\n");
    std::printf("energy %.17g\n", r.energy);
    std::printf("expected_energy %.17g\n", r.expected_energy);
    std::printf("prob %.17g\n", r.prob);
    std::printf("max_trunc %.17g\n", r.max_trunc);
    print_vec("entropy", r.entropy);
    std::printf("seconds %.6f\n", sec);
    return 0;
}

// The per-step loop of mps_anneal restated through the public MPS methods so
// that the bond dimension after every Trotter step can be observed.
static int mode_trace() {
    Inst I;
    if (!read_inst(I)) { std::fprintf(stderr, "bad instance\n"); return 2; }
    MPS psi(I.n);
    double trunc = 0.0;
    std::vector<double> bonds, ent, norms;
    auto half = [&](double A, double B) {
        for (int i = 0; i < I.n; ++i) {
            CMat g(2, 2);
            g(0, 0) = cd(-B * I.h[i], 0); g(1, 1) = cd(B * I.h[i], 0);
            g(0, 1) = cd(-A, 0); g(1, 0) = cd(-A, 0);
            CMat U = hermitian_expm(g, cd(0, -1) * (I.dt / 2.0));
            cd G[2][2] = {{U(0, 0), U(0, 1)}, {U(1, 0), U(1, 1)}};
            psi.apply_single(i, G);
        }
    };
    const int ns = I.ns < 1 ? 1 : I.ns;
    for (int k = 0; k < I.steps; ++k) {
        half(I.a[k], I.b[k]);
        for (int s = 0; s < ns; ++s) {
            if (s % 2 == 0)
                for (int e = 0; e < I.m; ++e)
                    psi.apply_zz(I.ei[e], I.ej[e], -I.b[k] * I.ew[e] * (I.dt / ns), I.chi, trunc);
            else
                for (int e = I.m; e-- > 0;)
                    psi.apply_zz(I.ei[e], I.ej[e], -I.b[k] * I.ew[e] * (I.dt / ns), I.chi, trunc);
        }
        half(I.a[k], I.b[k]);
        ent.push_back(psi.single_site_entropy(0));
        bonds.push_back(psi.max_bond());
        norms.push_back(psi.norm2());
    }
    print_vec("max_bond", bonds);
    print_vec("entropy", ent);
    print_vec("norm2", norms);
    std::printf("max_trunc %.17g\n", trunc);
    return 0;
}

static int mode_svd(int N, int reps) {
    std::mt19937_64 rng(42);
    std::normal_distribution<double> g(0.0, 1.0);
    // M = Q1 diag(exp(-k/40)) Q2 with Q1, Q2 from Gram-Schmidt of Gaussian matrices.
    auto unitary = [&](int n) {
        CMat Q(n, n);
        for (auto& x : Q.a) x = cd(g(rng), g(rng));
        for (int j = 0; j < n; ++j) {
            for (int p = 0; p < j; ++p) {
                cd d(0, 0);
                for (int i = 0; i < n; ++i) d += std::conj(Q(i, p)) * Q(i, j);
                for (int i = 0; i < n; ++i) Q(i, j) -= d * Q(i, p);
            }
            double nr = 0;
            for (int i = 0; i < n; ++i) nr += std::norm(Q(i, j));
            nr = std::sqrt(nr);
            for (int i = 0; i < n; ++i) Q(i, j) /= nr;
        }
        return Q;
    };
    CMat Q1 = unitary(N), Q2 = unitary(N), M(N, N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            cd acc(0, 0);
            for (int k = 0; k < N; ++k) acc += Q1(i, k) * std::exp(-k / 40.0) * std::conj(Q2(j, k));
            M(i, j) = acc;
        }
    double best = 1e300;
    std::vector<double> S;
    for (int r = 0; r < reps; ++r) {
        CMat U, Vh;
        auto t0 = std::chrono::steady_clock::now();
        svd(M, U, S, Vh);
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (sec < best) best = sec;
    }
    std::printf("N %d seconds %.6f\n", N, best);
    std::printf("S0 %.17g S_last %.17g\n", S.front(), S.back());
    return 0;
}

static void print_cmat(const char* key, const CMat& A) {
    std::printf("%s %zu %zu", key, A.rows, A.cols);
    for (const cd& z : A.a) std::printf(" %.17g %.17g", z.real(), z.imag());
    std::printf("\n");
}

static int mode_linalg() {
    const cd I(0, 1);
    // Matrices of the upstream known-answer tests (test_linalg.cpp:30-98) plus
    // a fixed 4x3 and 3x5 for the two svd branches.
    CMat H(3, 3);
    H(0, 0) = 2; H(1, 1) = -1; H(2, 2) = 0.5;
    H(0, 1) = cd(1, 2); H(1, 0) = std::conj(H(0, 1));
    H(0, 2) = cd(-0.5, 1); H(2, 0) = std::conj(H(0, 2));
    H(1, 2) = cd(0, -1); H(2, 1) = std::conj(H(1, 2));
    std::vector<double> w; CMat V;
    hermitian_eig(H, w, V);
    print_vec("eig_w", w);
    print_cmat("eig_V", V);
    CMat X(2, 2); X(0, 1) = 1; X(1, 0) = 1;
    print_cmat("expm_X", hermitian_expm(X, cd(0, -1) * (M_PI / 2)));
    CMat H2(3, 3);
    H2(0, 0) = 1; H2(1, 1) = 2; H2(2, 2) = -1;
    H2(0, 1) = cd(0.3, 0.7); H2(1, 0) = std::conj(H2(0, 1));
    H2(1, 2) = cd(-0.2, 0.5); H2(2, 1) = std::conj(H2(1, 2));
    print_cmat("expm_H2", hermitian_expm(H2, cd(0, -1) * 0.83));
    for (int shape = 0; shape < 2; ++shape) {
        int r = shape == 0 ? 4 : 3, c = shape == 0 ? 3 : 5;
        CMat A(r, c);
        for (int i = 0; i < r; ++i)
            for (int j = 0; j < c; ++j)
                A(i, j) = cd(std::sin(1.0 + i * 1.7 + j * 0.3), std::cos(0.5 * i - j * 1.1));
        CMat U, Vh; std::vector<double> S;
        svd(A, U, S, Vh);
        std::printf("svd_shape %d %d\n", r, c);
        print_vec("svd_S", S);
        print_cmat("svd_U", U);
        print_cmat("svd_Vh", Vh);
    }
    (void)I;
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: ref_harness anneal|trace|svd N REPS|linalg\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "anneal") return mode_anneal();
    if (mode == "trace") return mode_trace();
    if (mode == "svd") return mode_svd(argc > 2 ? std::atoi(argv[2]) : 64, argc > 3 ? std::atoi(argv[3]) : 1);
    if (mode == "linalg") return mode_linalg();
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
