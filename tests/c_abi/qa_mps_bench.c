/*
 * Compiled caller of the plain upstream entry point, standing in for the Rust
 * crate's `--features cuda` build (no cargo in this environment).
 *
 * It declares tessera_qa_anneal exactly as rust/src/ffi.rs:20-35 does (usize ->
 * size_t, *const f64 -> const double*, *mut i8 -> int8_t*), links the library the
 * way rust/build.rs:11-22 asks (-L<dir> -ltessera_qa), and walks the QA(mps)
 * column of rust/src/bin/bench.rs:41-56: smooth(800) schedule
 * (rust/src/schedule.rs:35-46), chi = 32, dense random instances n = 6, 8, 10, 12
 * with fields and couplings uniform in [-1, 1] and edge probability 0.5
 * (bench.rs:19-30; its StdRng stream comes from the rand crate, which is not
 * here, so this harness draws from its own SplitMix64 stream).
 *
 * Usage: qa_mps_bench [steps [max_n]] (defaults 800 and 12, as in bench.rs).
 * Output: one JSON object per instance (the instance, the returned energy and
 * spins, the brute-force ground energy). Exit status 0 when every call returned
 * a finite energy equal to the Ising energy of the spins it wrote.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

extern double tessera_qa_anneal(const double* h, size_t n, const size_t* ei, const size_t* ej, const double* ew,
                                size_t m, const double* schedule_a, const double* schedule_b, size_t steps,
                                size_t chi, int8_t* out_s, double* out_entropy);

static uint64_t rng_state = 7;
static double uniform(double lo, double hi) { /* SplitMix64, top 53 bits */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return lo + (hi - lo) * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
}

/* E(s) = -sum_e w s_i s_j - sum_i h_i s_i, edges first (rust/src/ising.rs:23-32) */
static double ising_energy(size_t n, const double* h, size_t m, const size_t* ei, const size_t* ej, const double* ew,
                           const int8_t* s) {
    double e = 0.0;
    for (size_t k = 0; k < m; ++k) e -= ew[k] * (double)s[ei[k]] * (double)s[ej[k]];
    for (size_t i = 0; i < n; ++i) e -= h[i] * (double)s[i];
    return e;
}

int main(int argc, char** argv) {
    size_t steps = argc > 1 ? (size_t)strtoul(argv[1], NULL, 10) : 800;
    const size_t max_n = argc > 2 ? (size_t)strtoul(argv[2], NULL, 10) : 12;
    const size_t chi = 32;
    if (steps < 2 || steps > 800) steps = 800;
    double a[800], b[800];
    const double half_pi = 1.5707963267948966;
    for (size_t k = 0; k < steps; ++k) {
        const double s = (double)k / (double)(steps - 1), c = cos(half_pi * s), sn = sin(half_pi * s);
        a[k] = c * c;
        b[k] = sn * sn;
    }
    const size_t sizes[4] = {6, 8, 10, 12};
    int ok = 1;
    for (int q = 0; q < 4; ++q) {
        const size_t n = sizes[q];
        if (n > max_n) break;
        double h[12], ew[66];
        size_t ei[66], ej[66], m = 0;
        int8_t spins[12];
        double entropy[800];
        for (size_t i = 0; i < n; ++i) h[i] = uniform(-1.0, 1.0);
        for (size_t i = 0; i < n; ++i)
            for (size_t k = i + 1; k < n; ++k)
                if (uniform(0.0, 1.0) < 0.5) {
                    ei[m] = i;
                    ej[m] = k;
                    ew[m] = uniform(-1.0, 1.0);
                    ++m;
                }
        const double e = tessera_qa_anneal(h, n, ei, ej, ew, m, a, b, steps, chi, spins, entropy);
        double exact = INFINITY;
        for (uint64_t x = 0; x < (1ull << n); ++x) {
            int8_t t[12];
            for (size_t i = 0; i < n; ++i) t[i] = ((x >> i) & 1) ? 1 : -1;
            const double v = ising_energy(n, h, m, ei, ej, ew, t);
            if (v < exact) exact = v;
        }
        const double es = ising_energy(n, h, m, ei, ej, ew, spins);
        if (!isfinite(e) || fabs(e - es) > 1e-12) ok = 0;
        double smax = 0.0;
        for (size_t k = 0; k < steps; ++k) smax = entropy[k] > smax ? entropy[k] : smax;
        printf("{\"n\": %zu, \"h\": [", n);
        for (size_t i = 0; i < n; ++i) printf("%s%.17g", i ? ", " : "", h[i]);
        printf("], \"edges\": [");
        for (size_t k = 0; k < m; ++k) printf("%s[%zu, %zu, %.17g]", k ? ", " : "", ei[k], ej[k], ew[k]);
        printf("], \"energy\": %.17g, \"spins\": [", e);
        for (size_t i = 0; i < n; ++i) printf("%s%d", i ? ", " : "", (int)spins[i]);
        printf("], \"exact\": %.17g, \"entropy_max\": %.17g}\n", exact, smax);
        fflush(stdout);
    }
    return ok ? 0 : 1;
}
