// State-vector oracle: rust/src/quantum.rs restated in C++ (same operation order;
// Rust does not contract FMAs and calls the platform libm for cos/sin/log2, as
// this file does). TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <cmath>

namespace oracle {

// ising.rs:23-32 in the bit convention of quantum.rs:45-56 (bit i set => +1).
static double config_energy(std::size_t x, std::size_t n, const std::vector<double>& h,
                            const std::vector<std::size_t>& ei, const std::vector<std::size_t>& ej,
                            const std::vector<double>& ew) {
    double e = 0.0;
    for (std::size_t k = 0; k < ew.size(); ++k) {
        const double si = ((x >> ei[k]) & 1) ? 1.0 : -1.0;
        const double sj = ((x >> ej[k]) & 1) ? 1.0 : -1.0;
        e -= ew[k] * si * sj;
    }
    for (std::size_t i = 0; i < n; ++i) {
        const double si = ((x >> i) & 1) ? 1.0 : -1.0;
        e -= h[i] * si;
    }
    return e;
}

// quantum.rs:178-187
static void phase(std::vector<double>& re, std::vector<double>& im, const std::vector<double>& diag,
                  double scale) {
    for (std::size_t x = 0; x < re.size(); ++x) {
        const double phi = scale * diag[x];
        const double c = std::cos(phi), s = std::sin(phi);
        const double nr = re[x] * c + im[x] * s;
        const double ni = im[x] * c - re[x] * s;
        re[x] = nr;
        im[x] = ni;
    }
}

// quantum.rs:190-207: qubits ascending; for x with bit clear and b = x|bit.
static void transverse(std::vector<double>& re, std::vector<double>& im, std::size_t n, double theta) {
    const double c = std::cos(theta), s = std::sin(theta);
    const std::size_t dim = re.size();
    for (std::size_t q = 0; q < n; ++q) {
        const std::size_t bit = (std::size_t)1 << q;
        for (std::size_t x = 0; x < dim; ++x) {
            if (x & bit) continue;
            const std::size_t b = x | bit;
            const double ra = re[x], ia = im[x], rb = re[b], ib = im[b];
            re[x] = c * ra - s * ib;
            im[x] = c * ia + s * rb;
            re[b] = c * rb - s * ia;
            im[b] = c * ib + s * ra;
        }
    }
}

// quantum.rs:138-175: qubit-0 reduced density matrix from pairs (x, x+1).
static double qubit0_entropy(const std::vector<double>& re, const std::vector<double>& im, std::size_t n) {
    if (n < 2) return 0.0;
    double r00 = 0.0, r11 = 0.0, r01r = 0.0, r01i = 0.0;
    for (std::size_t x = 0; x < re.size(); x += 2) {
        const std::size_t y = x + 1;
        r00 += re[x] * re[x] + im[x] * im[x];
        r11 += re[y] * re[y] + im[y] * im[y];
        r01r += re[x] * re[y] + im[x] * im[y];
        r01i += im[x] * re[y] - re[x] * im[y];
    }
    const double total = r00 + r11;
    if (total <= 0.0) return 0.0;
    r00 /= total;
    r11 /= total;
    const double off2 = (r01r * r01r + r01i * r01i) / (total * total);
    const double mean = (r00 + r11) / 2.0;
    const double half = (r00 - r11) / 2.0;
    const double rad = std::sqrt(half * half + off2);
    auto clamp01 = [](double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); };
    const double lam[2] = {clamp01(mean + rad), clamp01(mean - rad)};
    double ent = 0.0;
    for (double l : lam)
        if (l > 1e-12) ent -= l * std::log2(l);
    return ent;
}

// quantum.rs:115-124
static double mean_energy(const std::vector<double>& re, const std::vector<double>& im,
                          const std::vector<double>& diag) {
    double e = 0.0, total = 0.0;
    for (std::size_t x = 0; x < re.size(); ++x) {
        const double p = re[x] * re[x] + im[x] * im[x];
        e += p * diag[x];
        total += p;
    }
    return e / total;
}

// quantum.rs:63-112
SvResult sv_anneal(const std::vector<double>& h, const std::vector<std::size_t>& ei,
                   const std::vector<std::size_t>& ej, const std::vector<double>& ew,
                   const std::vector<double>& sa, const std::vector<double>& sb, double dt,
                   bool want_state) {
    const std::size_t n = h.size();
    const std::size_t dim = (std::size_t)1 << n;
    std::vector<double> diag(dim);
    for (std::size_t x = 0; x < dim; ++x) diag[x] = config_energy(x, n, h, ei, ej, ew);
    const double amp = 1.0 / std::sqrt((double)dim);
    std::vector<double> re(dim, amp), im(dim, 0.0);
    SvResult out;
    const std::size_t steps = sa.size();
    out.entropy.reserve(steps);
    out.energy_trace.reserve(steps);
    for (std::size_t k = 0; k < steps; ++k) {
        const double a = sa[k], b = sb[k];
        phase(re, im, diag, b * dt * 0.5);
        transverse(re, im, n, a * dt);
        phase(re, im, diag, b * dt * 0.5);
        out.entropy.push_back(qubit0_entropy(re, im, n));
        out.energy_trace.push_back(mean_energy(re, im, diag));
    }
    double total = 0.0, expected = 0.0, best_p = -1.0;
    std::size_t best = 0;
    for (std::size_t x = 0; x < dim; ++x) {
        const double p = re[x] * re[x] + im[x] * im[x];
        total += p;
        expected += p * diag[x];
        if (p > best_p) {
            best_p = p;
            best = x;
        }
    }
    out.spins.resize(n);
    for (std::size_t i = 0; i < n; ++i) out.spins[i] = ((best >> i) & 1) ? 1 : -1;
    out.energy = diag[best];
    out.expected_energy = expected / total;
    out.success_prob = best_p / total;
    if (want_state) {
        out.re = re;
        out.im = im;
    }
    return out;
}

}  // namespace oracle
