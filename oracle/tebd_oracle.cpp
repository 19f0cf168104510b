// TEBD oracle: upstream MPS engine restated for bit-for-bit parity.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Every loop below keeps the upstream summation order, accumulator start values
// and skip conditions, because the product of this file is compared bitwise with
// the upstream engine built from cuda/src/*.cpp (oracle/Makefile target `ref`).
#include "oracle.hpp"

#include <algorithm>
#include <cmath>

namespace oracle {

// ---------------------------------------------------------------------------
// linalg.cpp:7-54. Cyclic Jacobi: <=100 sweeps, stop when the squared upper
// off-diagonal sum drops below 1e-30 (absolute), skip |a_pq| < 1e-300.
int sym_jacobi(const std::vector<double>& a_in, std::size_t n, std::vector<double>& w,
               std::vector<double>& vecs, int max_sweeps) {
    std::vector<double> a(a_in);
    vecs.assign(n * n, 0.0);
    for (std::size_t d = 0; d < n; ++d) vecs[d * n + d] = 1.0;
    if (n == 1) {
        w.assign(1, a[0]);
        return 0;
    }
    int sweeps = 0;
    for (; sweeps < max_sweeps; ++sweeps) {
        double off = 0.0;
        for (std::size_t p = 0; p < n; ++p)
            for (std::size_t q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-30) break;
        for (std::size_t p = 0; p < n; ++p) {
            for (std::size_t q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[p * n + p], aqq = a[q * n + q];
                const double th = (aqq - app) / (2.0 * apq);
                const double sg = th >= 0 ? 1.0 : -1.0;
                const double t = sg / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0);
                const double s = t * c;
                double* col_p = &a[p];
                double* col_q = &a[q];
                for (std::size_t k = 0; k < n; ++k) {  // columns p, q
                    const double x = col_p[k * n], y = col_q[k * n];
                    col_p[k * n] = c * x - s * y;
                    col_q[k * n] = s * x + c * y;
                }
                double* row_p = &a[p * n];
                double* row_q = &a[q * n];
                for (std::size_t k = 0; k < n; ++k) {  // rows p, q
                    const double x = row_p[k], y = row_q[k];
                    row_p[k] = c * x - s * y;
                    row_q[k] = s * x + c * y;
                }
                for (std::size_t k = 0; k < n; ++k) {  // eigenvector columns
                    const double x = vecs[k * n + p], y = vecs[k * n + q];
                    vecs[k * n + p] = c * x - s * y;
                    vecs[k * n + q] = s * x + c * y;
                }
            }
        }
    }
    w.assign(n, 0.0);
    for (std::size_t d = 0; d < n; ++d) w[d] = a[d * n + d];
    return sweeps;
}

// linalg.cpp:56-105. Real 2n embedding [[R,-I],[I,R]], Jacobi, descending
// sort (std::sort, not stable), greedy complex Gram-Schmidt that drops vectors
// whose residual norm is below 1e-7.
void herm_eig(const Mat& h, std::vector<double>& evals, Mat& evecs) {
    const std::size_t n = h.r, m = 2 * n;
    std::vector<double> emb(m * m, 0.0);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            const double re = h.at(i, j).real(), im = h.at(i, j).imag();
            emb[i * m + j] = re;
            emb[(i + n) * m + (j + n)] = re;
            emb[i * m + (j + n)] = -im;
            emb[(i + n) * m + j] = im;
        }
    std::vector<double> w2, v2;
    sym_jacobi(emb, m, w2, v2);

    std::vector<std::size_t> order(m);
    for (std::size_t t = 0; t < m; ++t) order[t] = t;
    std::sort(order.begin(), order.end(),
              [&](std::size_t x, std::size_t y) { return w2[x] > w2[y]; });

    evals.assign(n, 0.0);
    evecs = Mat(n, n);
    std::size_t taken = 0;
    std::vector<cplx> vec(n);
    for (std::size_t t = 0; t < m && taken < n; ++t) {
        const std::size_t col = order[t];
        for (std::size_t i = 0; i < n; ++i) vec[i] = cplx(v2[i * m + col], v2[(i + n) * m + col]);
        for (std::size_t p = 0; p < taken; ++p) {
            cplx dot(0, 0);
            for (std::size_t i = 0; i < n; ++i) dot += std::conj(evecs.at(i, p)) * vec[i];
            for (std::size_t i = 0; i < n; ++i) vec[i] -= dot * evecs.at(i, p);
        }
        double nn = 0.0;
        for (std::size_t i = 0; i < n; ++i) nn += std::norm(vec[i]);
        nn = std::sqrt(nn);
        if (nn < 1e-7) continue;
        for (std::size_t i = 0; i < n; ++i) evecs.at(i, taken) = vec[i] / nn;
        evals[taken] = w2[col];
        ++taken;
    }
}

// linalg.cpp:107-172. Thin SVD through the smaller Gram matrix; S = sqrt(max(w,0));
// the other factor is A v / S, left zero when S <= 1e-300.
void gram_svd(const Mat& a, Mat& u, std::vector<double>& s, Mat& vh) {
    const std::size_t r = a.r, c = a.c, k = std::min(r, c);
    if (c <= r) {
        Mat g(c, c);
        for (std::size_t i = 0; i < c; ++i)
            for (std::size_t j = 0; j < c; ++j) {
                cplx acc(0, 0);
                for (std::size_t t = 0; t < r; ++t) acc += std::conj(a.at(t, i)) * a.at(t, j);
                g.at(i, j) = acc;
            }
        std::vector<double> w;
        Mat vecs;
        herm_eig(g, w, vecs);
        s.assign(k, 0.0);
        u = Mat(r, k);
        vh = Mat(k, c);
        for (std::size_t t = 0; t < k; ++t) {
            const double sv = std::sqrt(std::max(w[t], 0.0));
            s[t] = sv;
            for (std::size_t j = 0; j < c; ++j) vh.at(t, j) = std::conj(vecs.at(j, t));
            if (sv > 1e-300)
                for (std::size_t row = 0; row < r; ++row) {
                    cplx acc(0, 0);
                    for (std::size_t j = 0; j < c; ++j) acc += a.at(row, j) * vecs.at(j, t);
                    u.at(row, t) = acc / sv;
                }
        }
    } else {
        Mat g(r, r);
        for (std::size_t i = 0; i < r; ++i)
            for (std::size_t j = 0; j < r; ++j) {
                cplx acc(0, 0);
                for (std::size_t t = 0; t < c; ++t) acc += a.at(i, t) * std::conj(a.at(j, t));
                g.at(i, j) = acc;
            }
        std::vector<double> w;
        Mat vecs;
        herm_eig(g, w, vecs);
        s.assign(k, 0.0);
        u = Mat(r, k);
        vh = Mat(k, c);
        for (std::size_t t = 0; t < k; ++t) {
            const double sv = std::sqrt(std::max(w[t], 0.0));
            s[t] = sv;
            for (std::size_t row = 0; row < r; ++row) u.at(row, t) = vecs.at(row, t);
            if (sv > 1e-300)
                for (std::size_t j = 0; j < c; ++j) {
                    cplx acc(0, 0);
                    for (std::size_t row = 0; row < r; ++row) acc += std::conj(vecs.at(row, t)) * a.at(row, j);
                    vh.at(t, j) = acc / sv;
                }
        }
    }
}

// linalg.cpp:174-197. exp(coeff*H) = V diag(exp(coeff*w)) V^dagger.
Mat herm_expm(const Mat& h, cplx coeff) {
    std::vector<double> w;
    Mat vecs;
    herm_eig(h, w, vecs);
    const std::size_t n = h.r;
    std::vector<cplx> d(n);
    for (std::size_t k = 0; k < n; ++k) d[k] = std::exp(coeff * w[k]);
    Mat scaled(n, n);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t k = 0; k < n; ++k) scaled.at(i, k) = vecs.at(i, k) * d[k];
    Mat out(n, n);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) {
            cplx acc(0, 0);
            for (std::size_t k = 0; k < n; ++k) acc += scaled.at(i, k) * std::conj(vecs.at(j, k));
            out.at(i, j) = acc;
        }
    return out;
}

// One-sided (Hestenes) Jacobi on the columns of b (m x k, m >= k), cyclic pair
// order, complex 2x2 rotations built from the pair Gram [[alpha, gamma],
// [conj(gamma), beta]]; stops when every pair satisfies |gamma| <= 1e-15 *
// sqrt(alpha * beta). On return b holds U diag(s) (unsorted) and v the
// accumulated unitary.
static int hestenes_cols(Mat& b, Mat& v) {
    const std::size_t m = b.r, k = b.c;
    v = Mat(k, k);
    for (std::size_t d = 0; d < k; ++d) v.at(d, d) = 1.0;
    int sweep = 0;
    for (; sweep < 60; ++sweep) {
        bool rotated = false;
        for (std::size_t p = 0; p + 1 < k; ++p)
            for (std::size_t q = p + 1; q < k; ++q) {
                double alpha = 0.0, beta = 0.0;
                cplx gamma(0, 0);
                for (std::size_t i = 0; i < m; ++i) {
                    alpha += std::norm(b.at(i, p));
                    beta += std::norm(b.at(i, q));
                    gamma += std::conj(b.at(i, p)) * b.at(i, q);
                }
                const double g = std::abs(gamma);
                if (g == 0.0 || g <= (double)(m < 8 ? 8 : m) * 2.220446049250313e-16 * std::sqrt(alpha * beta))
                    continue;
                rotated = true;
                const cplx ph = std::conj(gamma) / g;  // multiplies column q
                const double zeta = (beta - alpha) / (2.0 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (std::size_t i = 0; i < m; ++i) {
                    const cplx x = b.at(i, p), y = b.at(i, q) * ph;
                    b.at(i, p) = c * x - sn * y;
                    b.at(i, q) = sn * x + c * y;
                }
                for (std::size_t i = 0; i < k; ++i) {
                    const cplx x = v.at(i, p), y = v.at(i, q) * ph;
                    v.at(i, p) = c * x - sn * y;
                    v.at(i, q) = sn * x + c * y;
                }
            }
        if (!rotated) break;
    }
    return sweep;
}

int hestenes_svd(const Mat& a, Mat& u, std::vector<double>& s, Mat& vh) {
    const bool tall = a.c <= a.r;
    Mat b = tall ? a : Mat(a.c, a.r);
    if (!tall)
        for (std::size_t i = 0; i < a.r; ++i)
            for (std::size_t j = 0; j < a.c; ++j) b.at(j, i) = std::conj(a.at(i, j));
    Mat v;
    const int sweeps = hestenes_cols(b, v);
    const std::size_t m = b.r, k = b.c;
    std::vector<double> nrm(k, 0.0);
    for (std::size_t j = 0; j < k; ++j) {
        double acc = 0.0;
        for (std::size_t i = 0; i < m; ++i) acc += std::norm(b.at(i, j));
        nrm[j] = std::sqrt(acc);
    }
    std::vector<std::size_t> order(k);
    for (std::size_t j = 0; j < k; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) { return nrm[x] > nrm[y]; });
    s.assign(k, 0.0);
    Mat ub(m, k), vb(k, k);
    for (std::size_t t = 0; t < k; ++t) {
        const std::size_t j = order[t];
        s[t] = nrm[j];
        for (std::size_t i = 0; i < m; ++i) ub.at(i, t) = nrm[j] > 0 ? b.at(i, j) / nrm[j] : cplx(0, 0);
        for (std::size_t i = 0; i < k; ++i) vb.at(i, t) = v.at(i, j);
    }
    if (tall) {  // a = ub s vb^dagger
        u = ub;
        vh = Mat(k, a.c);
        for (std::size_t t = 0; t < k; ++t)
            for (std::size_t j = 0; j < a.c; ++j) vh.at(t, j) = std::conj(vb.at(j, t));
    } else {  // a^dagger = ub s vb^dagger  =>  a = vb s ub^dagger
        u = vb;
        vh = Mat(k, a.c);
        for (std::size_t t = 0; t < k; ++t)
            for (std::size_t j = 0; j < a.c; ++j) vh.at(t, j) = std::conj(ub.at(j, t));
    }
    return sweeps;
}

// ---------------------------------------------------------------------------
// mps.cpp:9-17. |+>^n, every site (1,2,1) with amplitude 1/sqrt(2).
Chain::Chain(int n) : site((std::size_t)n) {
    const double amp = 1.0 / std::sqrt(2.0);
    for (auto& t : site) {
        t = Tensor(1, 1);
        t.at(0, 0, 0) = amp;
        t.at(0, 1, 0) = amp;
    }
}

int Chain::max_bond() const {  // mps.cpp:19-23
    int m = 1;
    for (const auto& t : site) m = std::max(m, t.dr);
    return m;
}

// mps.cpp:27-37
void Chain::gate1(int i, const cplx g[2][2]) {
    Tensor& t = site[(std::size_t)i];
    Tensor out(t.dl, t.dr);
    for (int a = 0; a < t.dl; ++a)
        for (int b = 0; b < t.dr; ++b) {
            const cplx x0 = t.at(a, 0, b), x1 = t.at(a, 1, b);
            out.at(a, 0, b) = g[0][0] * x0 + g[0][1] * x1;
            out.at(a, 1, b) = g[1][0] * x0 + g[1][1] * x1;
        }
    t = std::move(out);
}

// mps.cpp:41-100. theta[(a,r1),(r2,b)] = sum G[r1r2][s1s2] sum_c A(a,s1,c) C(c,s2,b),
// Gram SVD, keep = min(chi, k) minus trailing S < 1e-14, trunc = max(trunc,
// discarded/total), A <- U[:, :keep], C <- diag(S) Vh[:keep].
void Chain::gate2(int i, const cplx g[4][4], int chi, double& trunc) {
    Tensor& left = site[(std::size_t)i];
    Tensor& right = site[(std::size_t)i + 1];
    const int dl = left.dl, dr = right.dr, mid = left.dr;
    Mat th((std::size_t)dl * 2, (std::size_t)dr * 2);
    for (int a = 0; a < dl; ++a)
        for (int b = 0; b < dr; ++b) {
            cplx pair[2][2];
            for (int s1 = 0; s1 < 2; ++s1)
                for (int s2 = 0; s2 < 2; ++s2) {
                    cplx acc(0, 0);
                    for (int c = 0; c < mid; ++c) acc += left.at(a, s1, c) * right.at(c, s2, b);
                    pair[s1][s2] = acc;
                }
            for (int r1 = 0; r1 < 2; ++r1)
                for (int r2 = 0; r2 < 2; ++r2) {
                    cplx acc(0, 0);
                    for (int s1 = 0; s1 < 2; ++s1)
                        for (int s2 = 0; s2 < 2; ++s2) acc += g[r1 * 2 + r2][s1 * 2 + s2] * pair[s1][s2];
                    th.at((std::size_t)(a * 2 + r1), (std::size_t)(r2 * dr + b)) = acc;
                }
        }
    Mat u, vh;
    std::vector<double> sv;
    if (svd_mode == 1)
        hestenes_svd(th, u, sv, vh);
    else
        gram_svd(th, u, sv, vh);

    const int kfull = (int)sv.size();
    double total = 0.0;
    for (double x : sv) total += x * x;
    int keep = std::min(chi, kfull);
    while (keep > 1 && sv[(std::size_t)keep - 1] < 1e-14) --keep;
    int keep_lo = std::min(chi, kfull), keep_hi = keep_lo;
    while (keep_lo > 1 && sv[(std::size_t)keep_lo - 1] < 1e-13) --keep_lo;
    while (keep_hi > 1 && sv[(std::size_t)keep_hi - 1] < 1e-15) --keep_hi;
    if (keep_lo != keep_hi) ++ambiguous;
    double disc = 0.0;
    for (int t = keep; t < kfull; ++t) disc += sv[(std::size_t)t] * sv[(std::size_t)t];
    if (total > 0) trunc = std::max(trunc, disc / total);

    Tensor nl(dl, keep);
    for (int a = 0; a < dl; ++a)
        for (int s1 = 0; s1 < 2; ++s1)
            for (int k = 0; k < keep; ++k) nl.at(a, s1, k) = u.at((std::size_t)(a * 2 + s1), (std::size_t)k);
    Tensor nr(keep, dr);
    for (int k = 0; k < keep; ++k)
        for (int s2 = 0; s2 < 2; ++s2)
            for (int b = 0; b < dr; ++b)
                nr.at(k, s2, b) = sv[(std::size_t)k] * vh.at((std::size_t)k, (std::size_t)(s2 * dr + b));
    left = std::move(nl);
    right = std::move(nr);
}

// mps.cpp:105-124: SWAP permutation and exp(-i theta Z(x)Z).
static void make_swap(cplx g[4][4]) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) g[r][c] = cplx(0, 0);
    g[0][0] = 1;
    g[1][2] = 1;
    g[2][1] = 1;
    g[3][3] = 1;
}

static void make_zz(double theta, cplx g[4][4]) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) g[r][c] = cplx(0, 0);
    const cplx iu(0, 1);
    const double zzval[4] = {1.0, -1.0, -1.0, 1.0};
    for (int d = 0; d < 4; ++d) g[d][d] = std::exp(-iu * theta * zzval[d]);
}

// mps.cpp:126-145: swaps (k-1,k) for k = j..i+2, ZZ on (i,i+1), swaps back.
void Chain::zz_routed(int i, int j, double theta, int chi, double& trunc) {
    if (i > j) std::swap(i, j);
    if (i == j) return;
    cplx zz[4][4];
    make_zz(theta, zz);
    if (j == i + 1) {
        gate2(i, zz, chi, trunc);
        return;
    }
    cplx sw[4][4];
    make_swap(sw);
    for (int k = j; k > i + 1; --k) gate2(k - 1, sw, chi, trunc);
    gate2(i, zz, chi, trunc);
    for (int k = i + 1; k < j; ++k) gate2(k, sw, chi, trunc);
}

// mps.cpp:150-178. Left sweep E[b,b'] += (zf A(a,s,b) E[a,a']) conj(A(a',s,b')).
cplx Chain::zstring_raw(const std::vector<int>& zs) const {
    std::vector<cplx> env(1, cplx(1, 0));
    int ket_dim = 1, bra_dim = 1;
    for (int p = 0; p < n(); ++p) {
        const Tensor& t = site[(std::size_t)p];
        const bool withz = std::find(zs.begin(), zs.end(), p) != zs.end();
        const int nb = t.dr;
        std::vector<cplx> next((std::size_t)nb * nb, cplx(0, 0));
        for (int a = 0; a < ket_dim; ++a)
            for (int ap = 0; ap < bra_dim; ++ap) {
                const cplx e = env[(std::size_t)a * bra_dim + ap];
                if (e == cplx(0, 0)) continue;
                for (int s = 0; s < 2; ++s) {
                    const double zf = (withz && s == 1) ? -1.0 : 1.0;
                    for (int b = 0; b < nb; ++b) {
                        const cplx ket = zf * t.at(a, s, b) * e;
                        if (ket == cplx(0, 0)) continue;
                        for (int bp = 0; bp < nb; ++bp)
                            next[(std::size_t)b * nb + bp] += ket * std::conj(t.at(ap, s, bp));
                    }
                }
            }
        env.swap(next);
        ket_dim = nb;
        bra_dim = nb;
    }
    return env[0];
}

double Chain::norm2() const { return zstring_raw({}).real(); }  // mps.cpp:180

double Chain::z(int i) const {  // mps.cpp:182-186
    const double nn = norm2();
    if (nn <= 0) return 0.0;
    return zstring_raw({i}).real() / nn;
}

double Chain::zz(int i, int j) const {  // mps.cpp:188-194
    const double nn = norm2();
    if (nn <= 0) return 0.0;
    std::vector<int> zs = {i, j};
    std::sort(zs.begin(), zs.end());
    return zstring_raw(zs).real() / nn;
}

// mps.cpp:197-216. Row vector times A(:, s_i, :) along the chain; s = 0 for +1.
double Chain::prob(const std::vector<int8_t>& spins) const {
    std::vector<cplx> vec(1, cplx(1, 0));
    int len = 1;
    for (int p = 0; p < n(); ++p) {
        const Tensor& t = site[(std::size_t)p];
        const int s = spins[(std::size_t)p] > 0 ? 0 : 1;
        std::vector<cplx> next((std::size_t)t.dr, cplx(0, 0));
        for (int b = 0; b < t.dr; ++b) {
            cplx acc(0, 0);
            for (int a = 0; a < t.dl && a < len; ++a) acc += vec[(std::size_t)a] * t.at(a, s, b);
            next[(std::size_t)b] = acc;
        }
        vec.swap(next);
        len = t.dr;
    }
    const cplx amp = vec[0];
    const double nn = norm2();
    return nn > 0 ? std::norm(amp) / nn : 0.0;
}

// mps.cpp:218-291. 1-site density matrix from full left/right environments,
// closed-form 2x2 eigenvalues clamped to [0,1], entropy in bits (lambda > 1e-12).
double Chain::entropy(int i) const {
    std::vector<cplx> lenv(1, cplx(1, 0));
    int lk = 1, lb = 1;
    for (int p = 0; p < i; ++p) {
        const Tensor& t = site[(std::size_t)p];
        const int nb = t.dr;
        std::vector<cplx> next((std::size_t)nb * nb, cplx(0, 0));
        for (int a = 0; a < lk; ++a)
            for (int ap = 0; ap < lb; ++ap) {
                const cplx e = lenv[(std::size_t)a * lb + ap];
                if (e == cplx(0, 0)) continue;
                for (int s = 0; s < 2; ++s)
                    for (int b = 0; b < nb; ++b) {
                        const cplx ket = t.at(a, s, b) * e;
                        if (ket == cplx(0, 0)) continue;
                        for (int bp = 0; bp < nb; ++bp)
                            next[(std::size_t)b * nb + bp] += ket * std::conj(t.at(ap, s, bp));
                    }
            }
        lenv.swap(next);
        lk = nb;
        lb = nb;
    }
    std::vector<cplx> renv(1, cplx(1, 0));
    int rk = 1, rb = 1;
    for (int p = n() - 1; p > i; --p) {
        const Tensor& t = site[(std::size_t)p];
        const int nb = t.dl;
        std::vector<cplx> next((std::size_t)nb * nb, cplx(0, 0));
        for (int b = 0; b < rk; ++b)
            for (int bp = 0; bp < rb; ++bp) {
                const cplx e = renv[(std::size_t)b * rb + bp];
                if (e == cplx(0, 0)) continue;
                for (int s = 0; s < 2; ++s)
                    for (int a = 0; a < nb; ++a) {
                        const cplx ket = t.at(a, s, b) * e;
                        if (ket == cplx(0, 0)) continue;
                        for (int ap = 0; ap < nb; ++ap)
                            next[(std::size_t)a * nb + ap] += ket * std::conj(t.at(ap, s, bp));
                    }
            }
        renv.swap(next);
        rk = nb;
        rb = nb;
    }
    const Tensor& t = site[(std::size_t)i];
    cplx rho[2][2] = {{cplx(0, 0), cplx(0, 0)}, {cplx(0, 0), cplx(0, 0)}};
    for (int s = 0; s < 2; ++s)
        for (int sp = 0; sp < 2; ++sp) {
            cplx acc(0, 0);
            for (int a = 0; a < t.dl; ++a)
                for (int ap = 0; ap < t.dl; ++ap) {
                    const cplx l = lenv[(std::size_t)a * t.dl + ap];
                    if (l == cplx(0, 0)) continue;
                    for (int b = 0; b < t.dr; ++b)
                        for (int bp = 0; bp < t.dr; ++bp)
                            acc += l * t.at(a, s, b) * std::conj(t.at(ap, sp, bp)) *
                                   renv[(std::size_t)b * t.dr + bp];
                }
            rho[s][sp] = acc;
        }
    const double tr = (rho[0][0] + rho[1][1]).real();
    if (tr <= 0) return 0.0;
    const double d0 = rho[0][0].real() / tr, d1 = rho[1][1].real() / tr;
    const double off2 = std::norm(rho[0][1] / tr);
    const double mean = (d0 + d1) / 2.0;
    const double rad = std::sqrt(std::max(0.0, ((d0 - d1) / 2) * ((d0 - d1) / 2) + off2));
    const double lam[2] = {std::min(1.0, std::max(0.0, mean + rad)), std::min(1.0, std::max(0.0, mean - rad))};
    double ent = 0.0;
    for (double x : lam)
        if (x > 1e-12) ent -= x * std::log2(x);
    return ent;
}

std::vector<cplx> Chain::dense() const {
    // Contract left to right; index x accumulates site p's physical index in bit p.
    std::vector<cplx> cur(2);  // rows: partial basis index, cols: current bond
    int rows = 1, bond = 1;
    cur.assign(1, cplx(1, 0));
    for (int p = 0; p < n(); ++p) {
        const Tensor& t = site[(std::size_t)p];
        std::vector<cplx> next((std::size_t)rows * 2 * t.dr, cplx(0, 0));
        for (int x = 0; x < rows; ++x)
            for (int s = 0; s < 2; ++s)
                for (int b = 0; b < t.dr; ++b) {
                    cplx acc(0, 0);
                    for (int a = 0; a < bond; ++a) acc += cur[(std::size_t)x * bond + a] * t.at(a, s, b);
                    next[((std::size_t)(x + (s << p))) * t.dr + b] = acc;
                }
        cur.swap(next);
        rows *= 2;
        bond = t.dr;
    }
    return cur;  // bond == 1 at the end
}

// mps.cpp:295-381 (driver) with the half single-site layers built as in
// mps.cpp:313-322: g = [[-B h, -A], [-A, B h]], gate = exp(-i dt/2 g).
TebdResult tebd_anneal(const std::vector<double>& h, const std::vector<int>& ei,
                       const std::vector<int>& ej, const std::vector<double>& ew,
                       const std::vector<double>& sa, const std::vector<double>& sb,
                       double dt, int chi, int zz_substeps, bool want_norms, bool want_state,
                       int svd_mode) {
    const int n = (int)h.size();
    const int steps = (int)sa.size();
    Chain psi(n);
    psi.svd_mode = svd_mode;
    TebdResult res;
    res.entropy.assign((std::size_t)steps, 0.0);
    res.max_trunc = 0.0;

    auto half_layer = [&](double A, double B) {
        for (int i = 0; i < n; ++i) {
            Mat g(2, 2);
            g.at(0, 0) = cplx(-B * h[(std::size_t)i], 0);
            g.at(1, 1) = cplx(B * h[(std::size_t)i], 0);
            g.at(0, 1) = cplx(-A, 0);
            g.at(1, 0) = cplx(-A, 0);
            const Mat u = herm_expm(g, cplx(0, -1) * (dt / 2.0));
            const cplx gate[2][2] = {{u.at(0, 0), u.at(0, 1)}, {u.at(1, 0), u.at(1, 1)}};
            psi.gate1(i, gate);
        }
    };

    const int ns = zz_substeps < 1 ? 1 : zz_substeps;
    const double sub = dt / ns;
    for (int k = 0; k < steps; ++k) {
        const double A = sa[(std::size_t)k], B = sb[(std::size_t)k];
        half_layer(A, B);
        for (int s = 0; s < ns; ++s) {
            if (s % 2 == 0) {
                for (std::size_t e = 0; e < ew.size(); ++e)
                    psi.zz_routed(ei[e], ej[e], -B * ew[e] * sub, chi, res.max_trunc);
            } else {
                for (std::size_t e = ew.size(); e-- > 0;)
                    psi.zz_routed(ei[e], ej[e], -B * ew[e] * sub, chi, res.max_trunc);
            }
        }
        half_layer(A, B);
        res.entropy[(std::size_t)k] = psi.entropy(0);
        res.max_bond.push_back(psi.max_bond());
        res.ambiguous.push_back((int)psi.ambiguous);
        psi.ambiguous = 0;
        if (want_norms) res.norm2.push_back(psi.norm2());
    }

    double eexp = 0.0;  // mps.cpp:361-367
    for (std::size_t e = 0; e < ew.size(); ++e) eexp -= ew[e] * psi.zz(ei[e], ej[e]);
    for (int i = 0; i < n; ++i) eexp -= h[(std::size_t)i] * psi.z(i);
    res.expected_energy = eexp;

    res.spins.assign((std::size_t)n, 1);  // mps.cpp:369-371, 0 -> +1
    for (int i = 0; i < n; ++i) res.spins[(std::size_t)i] = (psi.z(i) >= 0.0) ? 1 : -1;

    double en = 0.0;  // mps.cpp:374-378
    for (std::size_t e = 0; e < ew.size(); ++e)
        en -= ew[e] * res.spins[(std::size_t)ei[e]] * res.spins[(std::size_t)ej[e]];
    for (int i = 0; i < n; ++i) en -= h[(std::size_t)i] * res.spins[(std::size_t)i];
    res.energy = en;
    res.prob = psi.prob(res.spins);
    if (want_state) res.state = psi.dense();
    return res;
}

}  // namespace oracle
