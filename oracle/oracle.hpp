// CPU oracle for the TESSERA quantum-annealing hot path. TEST INFRASTRUCTURE
// ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it
// as the checker. The product library (tessera_amd/) never links or calls it.
//
// It restates, in the upstream arithmetic order, so that results can be pinned
// bit-for-bit against the upstream engine compiled from its own sources:
//   - cuda/src/linalg.cpp  (Jacobi eigensolver, Hermitian eig, Gram SVD, expm)
//   - cuda/src/mps.cpp     (MPS state, TEBD gates, swap network, observables,
//                           anneal driver)
//   - cuda/src/ffi.cpp     (dt = 0.1, zz_substeps = 1 for the plain C ABI)
//   - rust/src/quantum.rs  (exact state-vector anneal)
// Floating-point order of every sum and product follows those files; see the
// per-function citations in tebd_oracle.cpp and sv_oracle.cpp.
#pragma once
#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

// Dense row-major complex matrix.
struct Mat {
    std::size_t r = 0, c = 0;
    std::vector<cplx> v;
    Mat() = default;
    Mat(std::size_t rows, std::size_t cols) : r(rows), c(cols), v(rows * cols, cplx(0, 0)) {}
    cplx& at(std::size_t i, std::size_t j) { return v[i * c + j]; }
    const cplx& at(std::size_t i, std::size_t j) const { return v[i * c + j]; }
};

// ---- dense numerics (cuda/src/linalg.cpp) ----
// Real symmetric eigenpairs by cyclic Jacobi (linalg.cpp:7-54). Returns the
// number of sweeps performed.
int sym_jacobi(const std::vector<double>& a_in, std::size_t n,
               std::vector<double>& w, std::vector<double>& vecs, int max_sweeps = 100);
void herm_eig(const Mat& h, std::vector<double>& evals, Mat& evecs);  // linalg.cpp:56-105
void gram_svd(const Mat& a, Mat& u, std::vector<double>& s, Mat& vh);  // linalg.cpp:107-172
Mat herm_expm(const Mat& h, cplx coeff);                               // linalg.cpp:174-197

// ---- MPS / TEBD (cuda/src/mps.cpp) ----
struct Tensor {  // shape (dl, 2, dr); element (a, s, b) at (a*2 + s)*dr + b
    int dl = 1, dr = 1;
    std::vector<cplx> v;
    Tensor() : v(2, cplx(0, 0)) {}
    Tensor(int l, int r) : dl(l), dr(r), v((std::size_t)l * 2 * r, cplx(0, 0)) {}
    cplx& at(int a, int s, int b) { return v[((std::size_t)a * 2 + s) * dr + b]; }
    cplx at(int a, int s, int b) const { return v[((std::size_t)a * 2 + s) * dr + b]; }
};

struct Chain {
    std::vector<Tensor> site;
    // 0: upstream Gram SVD (bit-exact with cuda/src/linalg.cpp:107-172).
    // 1: accurate one-sided Jacobi SVD with the same keep rule; this is the
    //    upstream update in exact arithmetic, the definition the GPU follows.
    int svd_mode = 0;
    // Updates whose kept rank sits at the rounding level of theta: the rank kept
    // with a cut of 1e-13 differs from the one with 1e-15 (diagnostic only; the
    // keep rule itself is the upstream 1e-14, mps.cpp:82).
    long long ambiguous = 0;
    explicit Chain(int n);
    int n() const { return (int)site.size(); }
    int max_bond() const;
    void gate1(int i, const cplx g[2][2]);
    void gate2(int i, const cplx g[4][4], int chi, double& trunc);
    void zz_routed(int i, int j, double theta, int chi, double& trunc);
    cplx zstring_raw(const std::vector<int>& zs) const;
    double norm2() const;
    double z(int i) const;
    double zz(int i, int j) const;
    double prob(const std::vector<int8_t>& spins) const;
    double entropy(int i) const;
    // Dense amplitudes psi[x], physical index of site i = bit i of x (test helper).
    std::vector<cplx> dense() const;
};

struct TebdResult {
    std::vector<int8_t> spins;
    double energy = 0, expected_energy = 0, prob = 0, max_trunc = 0;
    std::vector<double> entropy;
    std::vector<int> max_bond;   // after every step (observability only)
    std::vector<int> ambiguous;  // updates of each step with a rank at the rounding level
    std::vector<double> norm2;   // after every step, only when requested
    std::vector<cplx> state;     // dense final state, only when requested
};

TebdResult tebd_anneal(const std::vector<double>& h, const std::vector<int>& ei,
                       const std::vector<int>& ej, const std::vector<double>& ew,
                       const std::vector<double>& sa, const std::vector<double>& sb,
                       double dt, int chi, int zz_substeps,
                       bool want_norms = false, bool want_state = false, int svd_mode = 0);

// Accurate thin SVD by one-sided (Hestenes) Jacobi in FP64, singular values
// descending: a = u diag(s) vh. Returns the number of sweeps.
int hestenes_svd(const Mat& a, Mat& u, std::vector<double>& s, Mat& vh);

// ---- exact state vector (rust/src/quantum.rs) ----
struct SvResult {
    std::vector<int8_t> spins;
    double energy = 0, expected_energy = 0, success_prob = 0;
    std::vector<double> entropy, energy_trace;
    std::vector<double> re, im;  // final state, only when requested
};

SvResult sv_anneal(const std::vector<double>& h, const std::vector<std::size_t>& ei,
                   const std::vector<std::size_t>& ej, const std::vector<double>& ew,
                   const std::vector<double>& sa, const std::vector<double>& sb,
                   double dt, bool want_state = false);

}  // namespace oracle
