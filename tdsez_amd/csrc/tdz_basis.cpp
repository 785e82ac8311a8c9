// Host set-up of the B-spline discretisation: knots, quadrature, basis tables,
// 1D Gram bands, knot coarsening, Boehm prolongation and the symmetric
// eigen-solvers of the separable preconditioner.
//
// Reference behaviour followed:
//   open uniform knots with continuity p-1         src/core.cpp:1899
//   Gauss-Legendre rule per element, weight*detJac src/core.cpp:1793, src/tensorops.cpp:42
//   1D bands M, S, X and the CAP band W_d          src/tensorops.cpp:40-57, src/assembly/cap.cpp:52-88
//   CAP profiles (Manolopoulos; 1/m scale)         src/assembly/assembly.cpp:170-208, :1892-1907
//   knot coarsening rule, knot-insertion P         src/multigrid.cpp:25-93 (behaviour; own algorithms:
//                                                  run-length scan, Oslo algorithm, collocation check)
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

#include "tdz_internal.hpp"

namespace tdz {

// ---------------------------------------------------------------- tridiagonal QL
// Implicit QL with Wilkinson shifts on a symmetric tridiagonal matrix
// (diagonal d, sub-diagonal e[1..n-1]); Z accumulates the transformations
// (row-major n x n, columns become eigenvectors).
static void tql(int n, std::vector<double> &d, std::vector<double> &e, std::vector<double> &Z)
{
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-17 * dd) break;
            }
            if (m != l) {
                if (++iter > 200) throw std::runtime_error("tql: no convergence");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1, c = 1, p = 0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z[(size_t)k * n + i + 1];
                        Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * f;
                        Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

// Householder reduction of a symmetric matrix to tridiagonal form with the
// orthogonal transformation accumulated in A (row-major).
static void householder_tridiag(int n, std::vector<double> &A, std::vector<double> &d, std::vector<double> &e)
{
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0, scale = 0;
        if (l > 0) {
            for (int k = 0; k <= l; ++k) scale += std::fabs(a(i, k));
            if (scale == 0.0) {
                e[i] = a(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    a(i, k) /= scale;
                    h += a(i, k) * a(i, k);
                }
                double f = a(i, l);
                double g = f >= 0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                a(i, l) = f - g;
                f = 0;
                for (int j = 0; j <= l; ++j) {
                    a(j, i) = a(i, j) / h;
                    g = 0;
                    for (int k = 0; k <= j; ++k) g += a(j, k) * a(i, k);
                    for (int k = j + 1; k <= l; ++k) g += a(k, j) * a(i, k);
                    e[j] = g / h;
                    f += e[j] * a(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = a(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) a(j, k) -= (f * e[k] + g * a(i, k));
                }
            }
        } else {
            e[i] = a(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0;
                for (int k = 0; k <= l; ++k) g += a(i, k) * a(k, j);
                for (int k = 0; k <= l; ++k) a(k, j) -= g * a(k, i);
            }
        }
        d[i] = a(i, i);
        a(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) a(j, i) = a(i, j) = 0.0;
    }
}

void sym_eig(int n, std::vector<double> &A, std::vector<double> &lam)
{
    std::vector<double> e;
    householder_tridiag(n, A, lam, e);
    tql(n, lam, e, A);
}

void gauss_legendre(int Q, double *x, double *w)
{
    // Jacobi matrix of the Legendre recurrence: zero diagonal, b_k = k / sqrt(4k^2 - 1)
    std::vector<double> d(Q, 0.0), e(Q, 0.0), Z((size_t)Q * Q, 0.0);
    for (int k = 1; k < Q; ++k) e[k] = k / std::sqrt(4.0 * k * k - 1.0);
    for (int i = 0; i < Q; ++i) Z[(size_t)i * Q + i] = 1.0;
    tql(Q, d, e, Z);
    std::vector<int> order(Q);
    for (int i = 0; i < Q; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < Q; ++i) {
        const int k = order[i];
        x[i] = d[k];
        w[i] = 2.0 * Z[k] * Z[k];  // first component of the normalised eigenvector
    }
    // symmetrise: nodes come in +- pairs, the weights of a pair are equal
    for (int i = 0; i < Q / 2; ++i) {
        const double xm = 0.5 * (x[Q - 1 - i] - x[i]), wm = 0.5 * (w[i] + w[Q - 1 - i]);
        x[i] = -xm; x[Q - 1 - i] = xm;
        w[i] = w[Q - 1 - i] = wm;
    }
    if (Q % 2) x[Q / 2] = 0.0;
}

// Non-zero basis functions and their derivatives at u in span k (Piegl & Tiller,
// The NURBS Book, algorithm A2.3): ders[j*(p+1)+r], j = derivative order.
static void ders_basis(int k, double u, int p, const std::vector<double> &U, int nd, double *ders)
{
    std::vector<double> ndu((size_t)(p + 1) * (p + 1)), left(p + 1), right(p + 1), a(2 * (p + 1));
    auto N = [&](int i, int j) -> double & { return ndu[(size_t)i * (p + 1) + j]; };
    N(0, 0) = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[k + 1 - j];
        right[j] = U[k + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            N(j, r) = right[r + 1] + left[j - r];
            const double temp = N(r, j - 1) / N(j, r);
            N(r, j) = saved + right[r + 1] * temp;
            saved = left[j - r] * temp;
        }
        N(j, j) = saved;
    }
    for (int j = 0; j <= p; ++j) ders[j] = N(j, p);
    for (int r = 0; r <= p; ++r) {
        double *a0 = a.data(), *a1 = a.data() + (p + 1);
        int s1 = 0, s2 = 1;
        double *as[2] = {a0, a1};
        as[0][0] = 1.0;
        for (int kk = 1; kk <= nd; ++kk) {
            double dd = 0.0;
            const int rk = r - kk, pk = p - kk;
            if (r >= kk) {
                as[s2][0] = as[s1][0] / N(pk + 1, rk);
                dd = as[s2][0] * N(rk, pk);
            }
            const int j1 = rk >= -1 ? 1 : -rk;
            const int j2 = (r - 1 <= pk) ? kk - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                as[s2][j] = (as[s1][j] - as[s1][j - 1]) / N(pk + 1, rk + j);
                dd += as[s2][j] * N(rk + j, pk);
            }
            if (r <= pk) {
                as[s2][kk] = -as[s1][kk - 1] / N(pk + 1, r);
                dd += as[s2][kk] * N(r, pk);
            }
            ders[kk * (p + 1) + r] = dd;
            std::swap(s1, s2);
        }
    }
    double fac = p;
    for (int kk = 1; kk <= nd; ++kk) {
        for (int j = 0; j <= p; ++j) ders[kk * (p + 1) + j] *= fac;
        fac *= (p - kk);
    }
}

static double manolopoulos_profile(double xq, double xmax, double kmin)
{
    const double c = 2.62206, c3 = c * c * c;
    const double a = 1.0 - 16.0 / c3, b = (1.0 - 17.0 / c3) / (c * c), delta = 0.2;
    const double Emin = 0.5 * kmin * kmin;
    const double width = c / (2.0 * delta * kmin), rcap = xmax - width;
    const double ax = std::fabs(xq);
    if (ax <= rcap) return 0.0;
    double r = c * ((ax - rcap) / (xmax - rcap));
    if (ax >= xmax) r = c * (1.0 - 2.220446049250313e-16);
    return Emin * (a * r - b * r * r * r + 4.0 / ((c - r) * (c - r)) - 4.0 / ((c + r) * (c + r)));
}

static double cap_profile(const PhysParams &ph, double lo, double hi, double u)
{
    if (ph.cap_kind == 0) return 0.0;
    const double sdist = std::min(u - lo, hi - u);
    if (ph.cap_kind == 1) {
        const double width = ph.cap[1];
        if (!(sdist < width)) return 0.0;
        return ph.cap[0] * (width - sdist) * (width - sdist);
    }
    const double kmin = ph.cap[0], width = 2.62206 / (2.0 * 0.2 * kmin);
    if (!(sdist < width)) return 0.0;
    return manolopoulos_profile(std::min(width - sdist, width * (1.0 - 1e-10)), width, kmin) / ph.mass;
}

std::vector<double> uniform_knots(int E, double lo, double hi, int p)
{
    std::vector<double> t(E + 2 * p + 1);
    for (int i = 0; i <= p; ++i) {
        t[i] = lo;
        t[E + p + i] = hi;
    }
    for (int i = 1; i < E; ++i) t[p + i] = lo + (hi - lo) * i / E;
    return t;
}

void build_degenerate_axis(Axis1D &A)
{
    A = Axis1D();
    A.E = 1; A.p = 0; A.Q = 1; A.n = 1; A.active = false;
    A.X = {0.0}; A.W = {1.0}; A.cap = {0.0}; A.off = {0};
    A.B0 = {1.0}; A.B1 = {0.0};
    A.M = {1.0}; A.S = {0.0}; A.Xb = {0.0}; A.Wc = {0.0}; A.Vs = {0.0};
    A.isWall = {0};
}

void build_axis(Axis1D &A, const std::vector<double> &knots, int p, int Q, bool wall, const PhysParams &ph)
{
    A = Axis1D();
    A.active = true;
    A.p = p; A.Q = Q; A.t = knots; A.wall = wall;
    A.n = (int)knots.size() - p - 1;
    A.lo = knots.front(); A.hi = knots.back();
    std::vector<int> spans;
    for (int k = p; k < A.n; ++k)
        if (knots[k + 1] > knots[k]) spans.push_back(k);
    A.E = (int)spans.size();
    const int E = A.E, nen = p + 1, bw = 2 * p + 1;
    std::vector<double> gx(Q), gw(Q), ders(2 * nen);
    gauss_legendre(Q, gx.data(), gw.data());
    A.X.resize((size_t)E * Q); A.W.resize((size_t)E * Q); A.cap.resize((size_t)E * Q);
    A.off.resize(E);
    A.B0.assign((size_t)E * nen * Q, 0.0); A.B1 = A.B0;
    A.isWall.assign(A.n, 0);
    if (wall) { A.isWall[0] = 1; A.isWall[A.n - 1] = 1; }
    for (int e = 0; e < E; ++e) {
        const int k = spans[e];
        A.off[e] = k - p;
        const double u0 = knots[k], u1 = knots[k + 1], J = 0.5 * (u1 - u0);
        for (int q = 0; q < Q; ++q) {
            const double u = 0.5 * (u0 + u1) + J * gx[q];
            A.X[(size_t)e * Q + q] = u;
            A.W[(size_t)e * Q + q] = gw[q] * J;
            A.cap[(size_t)e * Q + q] = cap_profile(ph, A.lo, A.hi, u);
            ders_basis(k, u, p, knots, 1, ders.data());
            for (int a = 0; a < nen; ++a) {
                // wall functions are removed from the space: their tables are zero
                const bool w = A.isWall[A.off[e] + a];
                A.B0[((size_t)e * nen + a) * Q + q] = w ? 0.0 : ders[a];
                A.B1[((size_t)e * nen + a) * Q + q] = w ? 0.0 : ders[nen + a];
            }
        }
    }
    const size_t sz = (size_t)A.n * bw;
    A.M.assign(sz, 0.0); A.S.assign(sz, 0.0); A.Xb.assign(sz, 0.0); A.Wc.assign(sz, 0.0); A.Vs.assign(sz, 0.0);
    // the part of s V that is a sum of 1D terms (all of the harmonic potential)
    const double vsc = ph.pot_kind == 1 ? ph.s * 0.5 * ph.pot[0] : 0.0;
    for (int e = 0; e < E; ++e)
        for (int q = 0; q < Q; ++q) {
            const double w = A.W[(size_t)e * Q + q], u = A.X[(size_t)e * Q + q], cq = A.cap[(size_t)e * Q + q];
            for (int a = 0; a < nen; ++a) {
                const double va = A.B0[((size_t)e * nen + a) * Q + q], da = A.B1[((size_t)e * nen + a) * Q + q];
                for (int b = 0; b < nen; ++b) {
                    const double vb = A.B0[((size_t)e * nen + b) * Q + q], db = A.B1[((size_t)e * nen + b) * Q + q];
                    const size_t idx = (size_t)(A.off[e] + a) * bw + (b - a + p);
                    A.M[idx] += w * va * vb;
                    A.S[idx] += w * da * db;
                    A.Xb[idx] += w * u * va * vb;
                    A.Wc[idx] += w * cq * va * vb;
                    A.Vs[idx] += w * vsc * u * u * va * vb;
                }
            }
        }
}

// One multigrid coarsening of a knot vector: scan the runs of equal knots in order and
// remove every second element boundary (runs 1, 3, 5, ... counted from the left end) that
// is a simple knot; repeated knots (C0 interfaces) and both ends stay. Returns false when
// the scan removes nothing (the rule of src/multigrid.cpp:25-40).
bool coarsen_knots(const std::vector<double> &U, std::vector<double> &C)
{
    C.clear();
    C.reserve(U.size());
    bool removed = false;
    size_t run = 0;   // ordinal of the current run of equal knots
    for (size_t first = 0; first < U.size(); ++run) {
        size_t past = first;
        while (past < U.size() && U[past] == U[first]) ++past;
        const bool simple = past - first == 1, at_end = past == U.size();
        if (run % 2 == 1 && simple && !at_end) removed = true;
        else C.insert(C.end(), U.begin() + first, U.begin() + past);
        first = past;
    }
    return removed;
}

// Non-empty knot spans (elements) of a knot vector: distinct values minus one.
int knot_spans(const std::vector<double> &U)
{
    std::vector<double> v(U);
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return (int)v.size() - 1;
}

// Knot-insertion (refinement) matrix of the nested spaces S_p(C) in S_p(F), F a
// refinement of C: coarse B-spline j = sum_i R(i, j) fine B-spline i. Row i holds the
// discrete B-splines of the Oslo algorithm (Cohen, Lyche and Riesenfeld 1980): the
// blossoms of the coarse B-splines at the fine knots F[i+1..i+p], evaluated on the coarse
// span mu with C[mu] <= F[i] < C[mu+1] by p convex-combination sweeps
//   a_k(j) = (x_k - C[j]) / (C[j+k] - C[j]) a_{k-1}(j)
//          + (C[j+k+1] - x_k) / (C[j+k+1] - C[j+1]) a_{k-1}(j+1),   x_k = F[i+k],
// from a_0 = 1 at j = mu. Only the terms with a nonzero a_{k-1} are formed; their
// denominators span [C[mu], C[mu+1]] and are positive. Rows are sparse (columns
// mu-p..mu), O(p^2) per fine function.
std::vector<std::vector<std::pair<int, double>>> refinement_matrix(const std::vector<double> &C,
                                                                   const std::vector<double> &F, int p)
{
    const int nf = (int)F.size() - p - 1, nc = (int)C.size() - p - 1;
    if (nf < nc || F.front() != C.front() || F.back() != C.back())
        throw std::runtime_error("refinement_matrix: the fine knots do not refine the coarse ones");
    std::vector<std::vector<std::pair<int, double>>> rows(nf);
    std::vector<double> a(p + 1), b(p + 1);
    for (int i = 0; i < nf; ++i) {
        const int mu = (int)(std::upper_bound(C.begin(), C.end(), F[i]) - C.begin()) - 1;
        if (mu < p || mu >= nc) throw std::runtime_error("refinement_matrix: knot vectors are not open");
        // a[r] holds column mu - k + r after sweep k
        a[0] = 1.0;
        for (int k = 1; k <= p; ++k) {
            const double x = F[i + k];
            for (int r = 0; r <= k; ++r) {
                const int j = mu - k + r;
                double v = 0.0;
                if (r >= 1) v += (x - C[j]) / (C[j + k] - C[j]) * a[r - 1];              // a_{k-1}(j)
                if (r < k) v += (C[j + k + 1] - x) / (C[j + k + 1] - C[j + 1]) * a[r];  // a_{k-1}(j+1)
                b[r] = v;
            }
            std::swap(a, b);
        }
        for (int r = 0; r <= p; ++r)
            if (a[r] != 0.0) rows[i].push_back({mu - p + r, a[r]});
    }
    return rows;
}

// Greville abscissae: the mean of the p interior knots of each B-spline.
std::vector<double> greville(const std::vector<double> &U, int p)
{
    const int n = (int)U.size() - p - 1;
    std::vector<double> g(n);
    double window = std::accumulate(U.begin() + 1, U.begin() + 1 + p, 0.0);
    for (int i = 0; i < n; ++i) {
        g[i] = window / p;
        if (i + 1 < n) window += U[i + 1 + p] - U[i + 1];
    }
    return g;
}

// Values of the p+1 B-splines of degree p on knots U that are nonzero at x, the first
// one's index in first.
void basis_values(const std::vector<double> &U, int p, double x, int &first, double *vals)
{
    const int n = (int)U.size() - p - 1;
    int k = (int)(std::upper_bound(U.begin(), U.end(), x) - U.begin()) - 1;
    k = std::max(p, std::min(k, n - 1));
    std::vector<double> d(2 * (p + 1));
    ders_basis(k, x, p, U, 0, d.data());
    first = k - p;
    std::copy(d.begin(), d.begin() + p + 1, vals);
}

// Largest deviation from N_c,j(x) = sum_i R(i, j) N_f,i(x) over three points in every
// fine element (the two-scale identity of nested spline spaces, which determines R).
double refinement_error(const std::vector<double> &C, const std::vector<double> &F, int p,
                        const std::vector<std::vector<std::pair<int, double>>> &R)
{
    const int nc = (int)C.size() - p - 1;
    std::vector<double> vf(p + 1), vc(p + 1), sum(nc, 0.0);
    double err = 0.0;
    for (size_t k = 0; k + 1 < F.size(); ++k) {
        if (!(F[k + 1] > F[k])) continue;
        for (double s : {0.1127, 0.5, 0.8873}) {
            const double x = F[k] + s * (F[k + 1] - F[k]);
            int f0, c0;
            basis_values(F, p, x, f0, vf.data());
            basis_values(C, p, x, c0, vc.data());
            for (int a = 0; a <= p; ++a)
                for (auto &e : R[f0 + a]) sum[e.first] += e.second * vf[a];
            for (int j = std::max(0, c0 - p - 1); j < std::min(nc, c0 + 2 * p + 2); ++j) {
                const double want = (j >= c0 && j <= c0 + p) ? vc[j - c0] : 0.0;
                err = std::max(err, std::fabs(sum[j] - want));
                sum[j] = 0.0;
            }
        }
    }
    return err;
}

void gen_eig_bands(const Axis1D &A, const std::vector<double> &Tband, std::vector<double> &V,
                   std::vector<double> &lam)
{
    const int n = A.n, p = A.p, bw = 2 * p + 1;
    std::vector<int> act;
    for (int i = 0; i < n; ++i)
        if (!A.isWall[i]) act.push_back(i);
    const int m = (int)act.size();
    auto band = [&](const std::vector<double> &B, int i, int j) {
        return std::abs(i - j) > p ? 0.0 : B[(size_t)i * bw + (j - i + p)];
    };
    // M = L L^T (dense Cholesky of the active block)
    std::vector<double> L((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double s = band(A.M, act[j], act[j]);
        for (int k = std::max(0, j - p); k < j; ++k) s -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
        if (!(s > 0)) throw std::runtime_error("mass band is not positive definite");
        const double ljj = std::sqrt(s);
        L[(size_t)j * m + j] = ljj;
        for (int i = j + 1; i < std::min(m, j + p + 1); ++i) {
            double t = band(A.M, act[i], act[j]);
            for (int k = std::max(0, i - p); k < j; ++k) t -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
            L[(size_t)i * m + j] = t / ljj;
        }
    }
    // C = L^-1 T L^-T
    std::vector<double> Y((size_t)m * m, 0.0), Cm((size_t)m * m, 0.0);
    for (int c = 0; c < m; ++c)
        for (int i = 0; i < m; ++i) {
            double s = band(Tband, act[i], act[c]);
            for (int k = std::max(0, i - p); k < i; ++k) s -= L[(size_t)i * m + k] * Y[(size_t)k * m + c];
            Y[(size_t)i * m + c] = s / L[(size_t)i * m + i];
        }
    for (int r = 0; r < m; ++r)
        for (int i = 0; i < m; ++i) {
            double s = Y[(size_t)r * m + i];
            for (int k = std::max(0, i - p); k < i; ++k) s -= L[(size_t)i * m + k] * Cm[(size_t)r * m + k];
            Cm[(size_t)r * m + i] = s / L[(size_t)i * m + i];
        }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (Cm[(size_t)i * m + j] + Cm[(size_t)j * m + i]);
            Cm[(size_t)i * m + j] = Cm[(size_t)j * m + i] = v;
        }
    std::vector<double> lm;
    sym_eig(m, Cm, lm);
    // V = L^-T Q
    std::vector<double> Vm((size_t)m * m, 0.0);
    for (int c = 0; c < m; ++c)
        for (int i = m - 1; i >= 0; --i) {
            double s = Cm[(size_t)i * m + c];
            for (int k = i + 1; k < std::min(m, i + p + 1); ++k) s -= L[(size_t)k * m + i] * Vm[(size_t)k * m + c];
            Vm[(size_t)i * m + c] = s / L[(size_t)i * m + i];
        }
    V.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) V[(size_t)act[a] * n + act[b]] = Vm[(size_t)a * m + b];
    for (int b = 0; b < m; ++b) lam[act[b]] = lm[b];
}

}  // namespace tdz
