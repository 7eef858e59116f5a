// radiant_oracle.cpp -- CPU oracle for the discrete-ordinates (SN) transport solve.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain FP64 C++, the algorithm of
// Radiant.jl's `compute_flux(::SN)` path so that the HIP backend can be checked against it.
// Only the tests, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may
// load the library built from it; the product path never does.
//
// Every function cites the reference file:line it follows (paths relative to Radiant.jl's
// repository root).  Loop orders and the order of floating-point operations follow the
// reference where they influence the result; the compiler is asked not to contract
// multiply-adds (-ffp-contract=off) so that each product is rounded as in Julia.
//
// Scope: Cartesian 1D/2D/3D, solver types BTE, BFP, BCSD, FP, CSD and BFP-EF (slowing down by
// the DD scheme of order 1 in energy), diamond difference of order 1 on every axis
// (`scheme_weights.jl:47-51` with order 1 gives w = [-1, 2]) and, for the BTE, the DD / DG
// closures of higher orders (`sn_1D_bte.jl`, `sn_2D_bte.jl`, `sn_3D_bte.jl`), "standard", "galerkin-d" and
// "galerkin-m" angular operators, void / reflective / periodic boundaries, volume sources,
// accelerations none / livolant / anderson / gmres / bicgstab.
//
// Parity status of each part is recorded in the test-suite: quadratures, harmonics, M/D,
// the DD cell solve, the sweep, the Krylov solvers and the group loop are each pinned by a
// closed form, an invariant, a library routine or a brute-force dense solve.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>
#include <limits>
#include <map>
#include <array>
#ifdef _OPENMP
#include <omp.h>
#endif

using std::vector;
using std::string;

namespace {

thread_local string g_err;

[[noreturn]] void fail(const string& m) { throw std::runtime_error(m); }

const double PI = 3.141592653589793;  // Float64(pi)
const double EPS = std::numeric_limits<double>::epsilon();

// ---------------------------------------------------------------------------------------
// Quadratures
// ---------------------------------------------------------------------------------------

// tools/gauss_legendre.jl:18-74 -- Tricomi initial guess, Newton to eps on Bonnet's recursion.
void gauss_legendre(int N, vector<double>& x, vector<double>& w) {
    x.assign(N, 0.0);
    w.assign(N, 0.0);
    for (int i = 1; i <= N / 2; ++i) {
        double phi = ((PI / 2) * (double)(4 * i - 1)) / (double)(2 * N + 1);
        double s = std::sin(phi);
        double t = (1.0 - (double)(N - 1) / (double)(8 * N * N * N)) -
                   (1.0 / (double)(384LL * N * N * N * N)) * (39.0 - 28.0 / (s * s));
        x[i - 1] = t * std::cos(phi);
        x[N - i] = -x[i - 1];
    }
    double dPdx = 0.0;
    for (int i = 0; i < N; ++i) {
        double xold = 2.0;
        while (std::fabs(x[i] - xold) > EPS) {
            xold = x[i];
            double P1 = 1.0, P2 = x[i];
            for (int n = 2; n <= N; ++n) {
                double Pn = ((double)(2 * n - 1) * x[i] * P2 - (double)(n - 1) * P1) / (double)n;
                P1 = P2;
                P2 = Pn;
            }
            dPdx = (double)N * (P1 - x[i] * P2) / (1.0 - x[i] * x[i]);
            x[i] -= P2 / dPdx;
        }
        w[i] = 2.0 / ((1.0 - x[i] * x[i]) * (dPdx * dPdx));
    }
}

// tools/gauss_lobatto.jl:18-76 -- nodes +-1 and the roots of P'_{N-1}, by the same iteration from
// x = cos(pi i/(N-1)); weights 2/(N(N-1) P_{N-1}(x)^2); the centre node of an odd rule is 0.
void gauss_lobatto(int N, vector<double>& x, vector<double>& w) {
    if (N < 2) fail("Gauss-Lobatto needs at least 2 points");
    x.assign(N, 0.0);
    w.assign(N, 0.0);
    for (int i = 0; i < N; ++i) x[i] = std::cos(PI * (double)i / (double)(N - 1));
    w[0] = w[N - 1] = 2.0 / ((double)N * (double)(N - 1));
    for (int i = 1; i < N - 1; ++i) {
        double xold = 2.0, P2 = 0.0;
        for (int it = 0; it < 200 && std::fabs(x[i] - xold) > EPS; ++it) {
            xold = x[i];
            double P1 = 1.0;
            P2 = x[i];
            for (int n = 2; n < N; ++n) {
                double Pn = ((double)(2 * n - 1) * x[i] * P2 - (double)(n - 1) * P1) / (double)n;
                P1 = P2;
                P2 = Pn;
            }
            double d1 = (double)N * (P1 - x[i] * P2) / (1.0 - x[i] * x[i]);
            double d2 = (2.0 * x[i] * d1 - (double)N * (double)(N + 1) * P2) / (1.0 - x[i] * x[i]);
            x[i] -= d1 / d2;
        }
        w[i] = 2.0 / ((double)N * (double)(N - 1) * P2 * P2);
    }
    if (N % 2 == 1) x[(N + 1) / 2 - 1] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j)
            if (x[i] == x[j]) fail("Unable to build Lobatto quadrature at order N = " + std::to_string(N) + ".");
}

// tools/lebedev.jl for the orders whose rules are known in closed form (3, 5, 7, 9, 11).  Each
// rule is a list of orbits of the octahedral group O_h with one weight (fraction of the sphere):
// 1 = (1,0,0) 6 points, 2 = (0,a,a) a=1/sqrt2 12, 3 = (a,a,a) a=1/sqrt3 8, 4 = (a,a,b)
// b=sqrt(1-2a^2) 24, 5 = (a,b,0) b=sqrt(1-a^2) 24.  Weights are scaled to 4 pi.
static void oh_orbit(int code, double a, double v, vector<double>& mu, vector<double>& eta, vector<double>& xi,
                     vector<double>& w) {
    const int s2[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    auto push = [&](double x, double y, double z) {
        mu.push_back(x); eta.push_back(y); xi.push_back(z); w.push_back(4.0 * PI * v);
    };
    if (code == 1) {
        for (int ax = 0; ax < 3; ++ax)
            for (double sg : {1.0, -1.0}) {
                double p[3] = {0.0, 0.0, 0.0};
                p[ax] = sg;
                push(p[0], p[1], p[2]);
            }
    } else if (code == 2) {
        double c = 1.0 / std::sqrt(2.0);
        for (int z = 0; z < 3; ++z) {
            int o0 = z == 0 ? 1 : 0, o1 = z == 2 ? 1 : 2;
            for (auto& s : s2) {
                double p[3] = {0.0, 0.0, 0.0};
                p[o0] = s[0] * c; p[o1] = s[1] * c;
                push(p[0], p[1], p[2]);
            }
        }
    } else if (code == 3) {
        double c = 1.0 / std::sqrt(3.0);
        for (auto& s : s2)
            for (int s3 : {1, -1}) push(s[0] * c, s[1] * c, s3 * c);
    } else if (code == 4) {
        double b = std::sqrt(1.0 - 2.0 * a * a);
        for (int bi = 0; bi < 3; ++bi)
            for (auto& s : s2)
                for (int s3 : {1, -1}) {
                    int sg[3] = {s[0], s[1], s3};
                    double p[3] = {sg[0] * a, sg[1] * a, sg[2] * a};
                    p[bi] = sg[bi] * b;
                    push(p[0], p[1], p[2]);
                }
    } else if (code == 5) {
        double b = std::sqrt(1.0 - a * a);
        for (int z = 0; z < 3; ++z) {
            int o0 = z == 0 ? 1 : 0, o1 = z == 2 ? 1 : 2;
            const double pq[2][2] = {{a, b}, {b, a}};
            for (auto& t : pq)
                for (auto& s : s2) {
                    double p[3] = {0.0, 0.0, 0.0};
                    p[o0] = s[0] * t[0]; p[o1] = s[1] * t[1];
                    push(p[0], p[1], p[2]);
                }
        }
    }
}

void lebedev(int N, vector<double>& mu, vector<double>& eta, vector<double>& xi, vector<double>& w) {
    mu.clear(); eta.clear(); xi.clear(); w.clear();
    switch (N) {
        case 3: oh_orbit(1, 0.0, 1.0 / 6.0, mu, eta, xi, w); break;
        case 5:
            oh_orbit(1, 0.0, 1.0 / 15.0, mu, eta, xi, w);
            oh_orbit(3, 0.0, 3.0 / 40.0, mu, eta, xi, w);
            break;
        case 7:
            oh_orbit(1, 0.0, 1.0 / 21.0, mu, eta, xi, w);
            oh_orbit(2, 0.0, 4.0 / 105.0, mu, eta, xi, w);
            oh_orbit(3, 0.0, 9.0 / 280.0, mu, eta, xi, w);
            break;
        case 9:
            oh_orbit(1, 0.0, 1.0 / 105.0, mu, eta, xi, w);
            oh_orbit(3, 0.0, 9.0 / 280.0, mu, eta, xi, w);
            oh_orbit(5, std::sqrt((1.0 - 1.0 / std::sqrt(3.0)) / 2.0), 1.0 / 35.0, mu, eta, xi, w);
            break;
        case 11:
            oh_orbit(1, 0.0, 4.0 / 315.0, mu, eta, xi, w);
            oh_orbit(2, 0.0, 64.0 / 2835.0, mu, eta, xi, w);
            oh_orbit(3, 0.0, 27.0 / 1280.0, mu, eta, xi, w);
            oh_orbit(4, 1.0 / std::sqrt(11.0), 14641.0 / 725760.0, mu, eta, xi, w);
            break;
        default:
            fail("Lebedev quadrature is not available at order " + std::to_string(N) + " (available: 3, 5, 7, 9, 11).");
    }
}

// tools/gauss_legendre_chebychev.jl:19-39 -- product of GL (sorted ascending) in mu and a
// 2N-point Chebyshev rule in phi; polar axis is x.
void gauss_legendre_chebychev(int N, vector<double>& mu, vector<double>& eta, vector<double>& xi,
                              vector<double>& w) {
    vector<double> mun, wn;
    gauss_legendre(N, mun, wn);
    vector<int> idx(N);
    for (int i = 0; i < N; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return mun[a] < mun[b]; });
    vector<double> ms(N), ws(N);
    for (int i = 0; i < N; ++i) { ms[i] = mun[idx[i]]; ws[i] = wn[idx[i]]; }
    int Nd = 2 * N * N;
    mu.assign(Nd, 0); eta.assign(Nd, 0); xi.assign(Nd, 0); w.assign(Nd, 0);
    int i = 0;
    for (int n = 0; n < N; ++n)
        for (int m = 1; m <= 2 * N; ++m) {
            double phi = PI / (double)N * ((double)m - 0.5);
            mu[i] = ms[n];
            double st = std::sqrt(1.0 - ms[n] * ms[n]);
            eta[i] = st * std::cos(phi);
            xi[i] = st * std::sin(phi);
            w[i] = ws[n] * PI / (double)N;
            ++i;
        }
}

// tools/newton_bisection.jl:20-75
template <class F, class DF>
double newton_bisection(F f, DF dfdx, double x1, double x2, double eps0 = EPS, int Nmax = 100) {
    double f1 = f(x1), f2 = f(x2);
    if (f1 * f2 > 0.0) fail("newton_bisection: root not bracketed");
    if (f1 == 0.0) return x1;
    if (f2 == 0.0) return x2;
    double xl, xh;
    if (f1 < 0) { xl = x1; xh = x2; } else { xl = x2; xh = x1; }
    double rts = 0.5 * (x1 + x2);
    double dxold = std::fabs(x2 - x1);
    double dx = dxold;
    double frts = f(rts), dfrts = dfdx(rts);
    for (int j = 1; j <= Nmax; ++j) {
        if (((rts - xh) * dfrts - frts) * ((rts - xl) * dfrts - frts) > 0.0 ||
            std::fabs(2.0 * frts) > std::fabs(dxold * dfrts)) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
        } else {
            dxold = dx;
            dx = frts / dfrts;
            rts -= dx;
        }
        if (std::fabs(dx) < eps0) return rts;
        frts = f(rts);
        dfrts = dfdx(rts);
        if (frts < 0.0) xl = rts; else xh = rts;
    }
    fail("newton_bisection: maximum number of iterations exceeded");
}

// tools/carlson.jl:20-94 -- Carlson's equal-weight set, N(N+2) directions, octant order
// (+++, ++-, +-+, +--, -++, -+-, --+, ---).
void carlson(int N, vector<double>& mu, vector<double>& eta, vector<double>& xi, vector<double>& w) {
    if (N < 2 || N % 2 != 0) fail("Carlson quadrature available for even N >= 2.");
    int Nd = N * (N + 2), H = N / 2;
    mu.assign(Nd, 0); eta.assign(Nd, 0); xi.assign(Nd, 0); w.assign(Nd, 0);
    vector<double> wn(H), mum(H), mu0(H);
    for (int n = 1; n <= H; ++n) {
        wn[n - 1] = 4.0 * (double)(N - 2 * n + 2) / (double)(N * (N + 2));
        mum[n - 1] = 1.0 - (double)((N - 2 * n + 2) * (N - 2 * n + 4)) / (double)(N * (N + 2));
        mu0[n - 1] = 1.0 - (double)((N - 2 * n + 2) * (N - 2 * n + 2)) / (double)(N * (N + 2));
    }
    vector<double> mun(H);
    if (N == 2) {
        mun[0] = 1.0 / std::sqrt(3.0);
    } else {
        auto f = [&](double x) {
            double s = 0.0;
            for (int n = 0; n < H; ++n) { double t = mu0[n] + x * mum[n]; s += wn[n] * (t * t); }
            return 1.0 / 3.0 - s;
        };
        auto df = [&](double x) {
            double s = 0.0;
            for (int n = 0; n < H; ++n) s += 2.0 * wn[n] * mum[n] * (mu0[n] + x * mum[n]);
            return -s;
        };
        double x = newton_bisection(f, df, 1.0, -1.0);
        for (int n = 0; n < H; ++n) mun[n] = mu0[n] + x * mum[n];
    }
    double w0 = 1.0 / (double)(N * (N + 2));
    double A;
    if (N == 2) {
        A = 1.0;
    } else {
        auto g = [&](double a) {
            double s = 0.0;
            for (int n = 1; n <= H; ++n) {
                double m = mun[n - 1];
                s += (PI / 2) * wn[n - 1] * m -
                     (4 * PI) * w0 / std::sqrt(2.0) * std::sin(PI * a / 4) * std::sqrt(1.0 - m * m) /
                         std::sin(PI * a / (double)(2 * N - 4 * n + 4));
            }
            return s;
        };
        auto dg = [&](double a) {
            double s = 0.0;
            for (int n = 1; n <= H; ++n) {
                double m = mun[n - 1];
                double sq = std::sqrt(1.0 - m * m);
                double d = (double)(2 * N - 4 * n + 4);
                double sd = std::sin(PI * a / d);
                s += -4 * PI * w0 * std::sqrt(2.0) * PI / 8 *
                     (std::cos(PI * a / 4) * sq / sd -
                      2 * std::sin(PI * a / 4) * sq * std::cos(PI * a / d) / (sd * sd) / (double)(N - 2 * n + 2));
            }
            return s;
        };
        A = newton_bisection(g, dg, 0.5, 2.0);
    }
    static const int sgn[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                                  {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    int i = 0;
    for (int t = 0; t < 8; ++t)
        for (int n = 1; n <= H; ++n)
            for (int m = 1; m <= H - n + 1; ++m) {
                double phi = PI / 2 * ((double)(2 * m - 1) / (double)(N - 2 * n + 2) * A + (1.0 - A) / 2);
                mu[i] = sgn[t][0] * mun[n - 1];
                double st = std::sqrt(1.0 - mu[i] * mu[i]);
                eta[i] = sgn[t][1] * st * std::cos(phi);
                xi[i] = sgn[t][2] * st * std::sin(phi);
                w[i] = (4 * PI) * w0;
                ++i;
            }
}

void fold_upper(vector<double>& mu, vector<double>& eta, vector<double>& xi, vector<double>& w);

// tools/quadrature.jl:19-59 (gauss-legendre, gauss-lobatto, gauss-legendre-chebychev, lebedev,
// carlson; the Qdims=2 fold onto the upper hemisphere is at :36-54).
void quadrature(int N, const string& type, int Ndims, int Qdims, vector<double>& mu,
                vector<double>& eta, vector<double>& xi, vector<double>& w) {
    if (type == "gauss-legendre" && Ndims == 1) {
        gauss_legendre(N, mu, w);
        eta.assign(mu.size(), 0.0);   // sn_flux.jl:45 expands a 1D set to [mu, 0, 0]
        xi.assign(mu.size(), 0.0);
    } else if (type == "gauss-lobatto" && Ndims == 1) {
        gauss_lobatto(N, mu, w);
        eta.assign(mu.size(), 0.0);
        xi.assign(mu.size(), 0.0);
    } else if (type == "lebedev") {
        lebedev(N, mu, eta, xi, w);
    } else if (type == "gauss-legendre-chebychev") {
        gauss_legendre_chebychev(N, mu, eta, xi, w);
    } else if (type == "carlson") {
        carlson(N, mu, eta, xi, w);
    } else {
        fail("Unknown " + type + " quadrature in " + std::to_string(Ndims) + "D geometry.");
    }
    if (Qdims == 2) fold_upper(mu, eta, xi, w);
}

// quadrature.jl:36-54: the unit-sphere set folded onto the upper hemisphere (xi >= 0), each
// lower point's weight added to its mirror image.
void fold_upper(vector<double>& mu, vector<double>& eta, vector<double>& xi, vector<double>& w) {
    {
        int Nd = (int)w.size();
        vector<double> w0 = w;
        for (int i = 0; i < Nd; ++i) {
            if (xi[i] >= 0) continue;
            int best = 0;
            double bv = 0;
            for (int j = 0; j < Nd; ++j) {
                double a = std::fabs(mu[i] - mu[j]), b = std::fabs(eta[i] - eta[j]), c = std::fabs(xi[i] + xi[j]);
                double e = a * a + b * b + c * c;
                if (j == 0 || e < bv) { bv = e; best = j; }
            }
            w0[best] += w0[i];
        }
        vector<double> m2, e2, x2, ww;
        for (int i = 0; i < Nd; ++i)
            if (xi[i] >= 0) { m2.push_back(mu[i]); e2.push_back(eta[i]); x2.push_back(xi[i]); ww.push_back(w0[i]); }
        mu = m2; eta = e2; xi = x2; w = ww;
    }
}

// ---------------------------------------------------------------------------------------
// Polynomials and harmonics
// ---------------------------------------------------------------------------------------

// tools/factorial_factor.jl:17-33 -- exp(sum log) form of factorial ratios.
double factorial_factor(const vector<int>& num, const vector<int>& den,
                        const vector<std::pair<double, double>>& ex = {}) {
    double f = 0.0;
    for (int n : num) for (int i = 1; i <= n; ++i) f += std::log((double)i);
    for (int n : den) for (int i = 1; i <= n; ++i) f -= std::log((double)i);
    for (auto& e : ex) {
        double x = e.first, m = e.second;
        if (m == 0.0) continue;
        if (x == 0.0) return m > 0 ? 0.0 : INFINITY;
        f += m * std::log(x);
    }
    return std::exp(f);
}

// tools/legendre_polynomials.jl:52-75 -- Bonnet recursion.
vector<double> legendre_up_to_L(int L, double x) {
    if (L < 0) fail("Legendre order is greater or equal to zero.");
    vector<double> P(L + 1, 0.0);
    if (x == 1.0) {
        for (int l = 0; l <= L; ++l) P[l] = 1.0;
    } else if (x == -1.0) {
        for (int l = 0; l <= L; ++l) P[l] = (l % 2 == 0) ? 1.0 : -1.0;
    } else {
        P[0] = 1.0;
        if (L >= 1) P[1] = x;
        for (int l = 2; l <= L; ++l)
            P[l] = ((double)(2 * l - 1) * x * P[l - 1] - (double)(l - 1) * P[l - 2]) / (double)l;
    }
    return P;
}

// tools/legendre_polynomials.jl:91-93
vector<double> half_range_legendre_up_to_L(int L, double x) {
    vector<double> P = legendre_up_to_L(L, 2 * x - 1);
    for (int l = 0; l <= L; ++l) P[l] = std::sqrt((double)(2 * l + 1)) * P[l];
    return P;
}

// tools/jacobi_polynomials.jl:16-49 -- three-term recurrence (Darboux asymptotics at l>=125).
vector<double> jacobi_up_to_L(int L, int a, int b, double x) {
    vector<double> P(L + 1, 0.0);
    if (x == 1.0) {
        for (int l = 0; l <= L; ++l) P[l] = factorial_factor({l + a}, {l, a});
    } else if (x == -1.0) {
        for (int l = 0; l <= L; ++l) P[l] = ((l % 2 == 0) ? 1.0 : -1.0) * factorial_factor({l + b}, {l, b});
    } else {
        P[0] = 1.0;
        if (L >= 1) P[1] = (double)(a + 1) + 0.5 * (double)(a + b + 2) * (x - 1.0);
        for (int l = 2; l <= L; ++l) {
            if (l < 125) {
                long s = 2 * l + a + b;
                double inner = (double)(s * (s - 2)) * x + (double)(a * a) - (double)(b * b);
                double t1 = (double)(s - 1) * inner * P[l - 1];
                double t2 = (double)(2L * (l + a - 1) * (l + b - 1) * s) * P[l - 2];
                P[l] = (t1 - t2) / (double)(2L * l * (l + a + b) * (s - 2));
            } else {
                double th = std::acos(x);
                double kth = 1.0 / (std::sqrt(PI) * std::pow(std::sin(0.5 * th), a + 0.5) *
                                    std::pow(std::cos(0.5 * th), b + 0.5));
                double Nn = l + 0.5 * (a + b + 1);
                double gam = -0.5 * PI * (a + 0.5);
                P[l] = 1.0 / std::sqrt((double)l) * kth * std::cos(Nn * th + gam);
            }
        }
    }
    return P;
}

typedef vector<vector<double>> Harm;  // H[l][l+m], m = -l..l

// tools/real_spherical_harmonics.jl:52-88 -- R_lm = C_lm P_l^|m|(mu) T_m(phi), no
// Condon-Shortley phase, T_m = cos(m phi) for m>=0 and sin(|m| phi) for m<0.
Harm real_sh_up_to_L(int L, double mu, double phi) {
    if (L < 0) fail("Legendre order is greater or equal to zero.");
    vector<vector<double>> Plm(L + 1);
    for (int l = 0; l <= L; ++l) Plm[l].assign(l + 1, 0.0);
    if (mu == -1.0 || mu == 1.0) {
        vector<double> P = jacobi_up_to_L(L, 0, 0, mu);
        for (int l = 0; l <= L; ++l) Plm[l][0] = P[l];
    } else {
        for (int m = 0; m <= L; ++m) {
            vector<double> P = jacobi_up_to_L(L - m, m, m, mu);
            for (int l = m; l <= L; ++l)
                Plm[l][m] = factorial_factor({l + m}, {l}, {{2.0, (double)(-m)}, {1.0 - mu * mu, m / 2.0}}) * P[l - m];
        }
    }
    Harm R(L + 1);
    for (int l = 0; l <= L; ++l) {
        R[l].assign(2 * l + 1, 0.0);
        for (int m = -l; m <= l; ++m) {
            int am = std::abs(m);
            double T = (m >= 0) ? std::cos((double)m * phi) : std::sin((double)am * phi);
            double C = std::sqrt((double)(2 - (m == 0)) * factorial_factor({l - am}, {l + am}));
            R[l][l + m] = C * Plm[l][am] * T;
        }
    }
    return R;
}

// tools/real_spherical_harmonics.jl:104-139 -- half-range harmonics on mu in [0,1].
Harm real_half_range_sh_up_to_L(int L, double mu, double phi) {
    if (L < 0) fail("Legendre order is greater or equal to zero.");
    if (!(0 <= mu && mu <= 1)) fail("Invalid direction cosine (should be between 0 and 1).");
    vector<vector<double>> Plm(L + 1);
    for (int l = 0; l <= L; ++l) Plm[l].assign(l + 1, 0.0);
    if (mu == -1.0 || mu == 1.0) {
        vector<double> P = jacobi_up_to_L(L, 0, 0, 2 * mu - 1);
        for (int l = 0; l <= L; ++l) Plm[l][0] = P[l];
    } else {
        double x = 2 * mu - 1;
        for (int m = 0; m <= L; ++m) {
            vector<double> P = jacobi_up_to_L(L - m, m, m, x);
            for (int l = m; l <= L; ++l)
                Plm[l][m] = factorial_factor({l + m}, {l}, {{2.0, (double)(-m)}, {1.0 - x * x, m / 2.0}}) * P[l - m];
        }
    }
    Harm R(L + 1);
    for (int l = 0; l <= L; ++l) {
        R[l].assign(2 * l + 1, 0.0);
        for (int m = -l; m <= l; ++m) {
            int am = std::abs(m);
            double T = (m >= 0) ? std::cos((double)m * phi) : std::sin((double)am * phi);
            double C = std::sqrt((double)(2 - (m == 0)) / (2 * PI) * (double)(2 * l + 1) *
                                 factorial_factor({l - am}, {l + am}));
            R[l][l + m] = C * Plm[l][am] * T;
        }
    }
    return R;
}

// ---------------------------------------------------------------------------------------
// Dense inverse (LU with partial pivoting), the role of LinearAlgebra.inv in the reference.
// Column-major n x n.
// ---------------------------------------------------------------------------------------
vector<double> invert(const vector<double>& A, int n) {
    vector<double> LU = A;
    vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double mx = std::fabs(LU[k + n * k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU[i + n * k]) > mx) { mx = std::fabs(LU[i + n * k]); p = i; }
        piv[k] = p;
        if (LU[p + n * k] == 0.0) fail("inv: singular matrix");
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(LU[k + n * j], LU[p + n * j]);
        for (int i = k + 1; i < n; ++i) {
            LU[i + n * k] /= LU[k + n * k];
            double l = LU[i + n * k];
            for (int j = k + 1; j < n; ++j) LU[i + n * j] -= l * LU[k + n * j];
        }
    }
    vector<double> X(n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        vector<double> b(n, 0.0);
        b[c] = 1.0;
        for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) b[i] -= LU[i + n * j] * b[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) b[i] -= LU[i + n * j] * b[j];
            b[i] /= LU[i + n * i];
        }
        for (int i = 0; i < n; ++i) X[i + n * c] = b[i];
    }
    return X;
}

// ---------------------------------------------------------------------------------------
// Angular operators
// ---------------------------------------------------------------------------------------

// particle_transport/angular_polynomial_basis.jl:9-17
int galerkin_minimum_order(int Nd, int Qdims) {
    if (Qdims == 1) return Nd - 1;
    int n = 0;
    while (true) {
        int count = (Qdims == 2) ? (n + 1) * (n + 2) / 2 : (n + 1) * (n + 1);
        if (count >= Nd) return n;
        ++n;
    }
}

struct Basis {
    int Np = 0;
    vector<double> M;  // Nd x Np column-major (M[n + Nd*p])
    vector<double> D;  // Np x Nd column-major (D[p + Np*n])
    vector<long> pl, pm;
};

// particle_transport/angular_polynomial_basis.jl:465-548 -- classical Gram-Schmidt selection of
// independent harmonics (threshold 1e-5 on the residual norm), then M = inv(D) (g_type 1) or
// D = inv(M) (g_type 2, 3).  Returns false when fewer than Nd harmonics were found.
bool gram_schmidt(int Nd, int L, const vector<Harm>& R, const vector<double>& w, int Qdims, int gtype,
                  Basis& B) {
    B.Np = Nd;
    B.pl.assign(Nd, 0);
    B.pm.assign(Nd, 0);
    B.M.assign((size_t)Nd * Nd, 0.0);
    B.D.assign((size_t)Nd * Nd, 0.0);
    vector<double> u((size_t)Nd * Nd, 0.0);  // u[n + Nd*k]
    auto nrm = [&](const vector<double>& a) {
        double s = 0;
        for (double v : a) s += v * v;
        return std::sqrt(s);
    };
    {
        vector<double> v(Nd);
        for (int n = 0; n < Nd; ++n) v[n] = R[n][0][0];
        double nv = nrm(v);
        for (int n = 0; n < Nd; ++n) {
            if (gtype == 1) B.D[0 + (size_t)Nd * n] = w[n] * R[n][0][0];
            else if (gtype == 2) B.M[n + 0] = 1 / (4 * PI) * R[n][0][0];
            else B.M[n + 0] = R[n][0][0];
            u[n] = R[n][0][0] / nv;
        }
    }
    int i = 1;  // number of kept harmonics
    for (int l = 1; l <= L && i < Nd; ++l) {
        int mmin = (Qdims == 2) ? 0 : -l;
        for (int m = mmin; m <= l; ++m) {
            vector<double> v(Nd), S(Nd, 0.0), vk(Nd);
            for (int n2 = 0; n2 < Nd; ++n2) v[n2] = R[n2][l][m + l];
            for (int n = 0; n < i; ++n) {
                double ip = 0;
                for (int k = 0; k < Nd; ++k) ip += u[k + (size_t)Nd * n] * v[k];
                for (int n2 = 0; n2 < Nd; ++n2) S[n2] += ip * u[n2 + (size_t)Nd * n];
            }
            for (int n = 0; n < Nd; ++n) vk[n] = R[n][l][m + l] - S[n];
            double nk = nrm(vk);
            if (nk > 1e-5) {
                for (int n = 0; n < Nd; ++n) {
                    u[n + (size_t)Nd * i] = vk[n] / nk;
                    if (gtype == 1) B.D[i + (size_t)Nd * n] = w[n] * R[n][l][m + l];
                    else if (gtype == 2) B.M[n + (size_t)Nd * i] = (double)(2 * l + 1) / (4 * PI) * R[n][l][m + l];
                    else B.M[n + (size_t)Nd * i] = R[n][l][m + l];
                }
                B.pl[i] = l;
                B.pm[i] = m;
                ++i;
            }
            if (i == Nd) break;
        }
    }
    if (i < Nd) return false;
    if (gtype == 1) B.M = invert(B.D, Nd);
    else B.D = invert(B.M, Nd);
    return true;
}

// particle_transport/angular_polynomial_basis.jl:51-179 (volume basis).
Basis angular_polynomial_basis(const vector<double>& mu, const vector<double>& eta, const vector<double>& xi,
                               const vector<double>& w, int L, const string& type, int Qdims) {
    int Nd = (int)w.size();
    bool gal = (type == "galerkin-m" || type == "galerkin-d");
    int Lb = gal ? std::max(L, galerkin_minimum_order(Nd, Qdims)) : L;
    Basis B;
    if (Qdims == 1) {
        vector<vector<double>> Pl(Nd);
        for (int n = 0; n < Nd; ++n) Pl[n] = legendre_up_to_L(Lb, mu[n]);
        if (type == "standard") {
            B.Np = L + 1;
            B.pl.assign(B.Np, 0); B.pm.assign(B.Np, 0);
            B.M.assign((size_t)Nd * B.Np, 0); B.D.assign((size_t)B.Np * Nd, 0);
            for (int l = 0; l <= L; ++l) {
                B.pl[l] = l;
                for (int n = 0; n < Nd; ++n) {
                    B.M[n + (size_t)Nd * l] = (double)(2 * l + 1) / 2 * Pl[n][l];
                    B.D[l + (size_t)B.Np * n] = w[n] * Pl[n][l];
                }
            }
        } else if (type == "galerkin-m") {
            B.Np = Nd;
            B.pl.assign(Nd, 0); B.pm.assign(Nd, 0);
            B.M.assign((size_t)Nd * Nd, 0);
            for (int l = 0; l < Nd; ++l) {
                B.pl[l] = l;
                for (int n = 0; n < Nd; ++n) B.M[n + (size_t)Nd * l] = (double)(2 * l + 1) / 2 * Pl[n][l];
            }
            B.D = invert(B.M, Nd);
        } else if (type == "galerkin-d") {
            B.Np = Nd;
            B.pl.assign(Nd, 0); B.pm.assign(Nd, 0);
            B.D.assign((size_t)Nd * Nd, 0);
            for (int l = 0; l < Nd; ++l) {
                B.pl[l] = l;
                for (int n = 0; n < Nd; ++n) B.D[l + (size_t)Nd * n] = w[n] * Pl[n][l];
            }
            B.M = invert(B.D, Nd);
        } else {
            fail("Unknown method.");
        }
    } else if (Qdims <= 3) {
        vector<double> phi(Nd);
        for (int n = 0; n < Nd; ++n) phi[n] = std::atan2(xi[n], eta[n]);
        vector<Harm> R(Nd);
        for (int n = 0; n < Nd; ++n) R[n] = real_sh_up_to_L(Lb, mu[n], phi[n]);
        if (type == "standard") {
            B.Np = (Qdims == 2) ? (L + 1) * (L + 2) / 2 : (L + 1) * (L + 1);
            B.pl.assign(B.Np, 0); B.pm.assign(B.Np, 0);
            B.M.assign((size_t)Nd * B.Np, 0); B.D.assign((size_t)B.Np * Nd, 0);
            int li = 0;
            for (int l = 0; l <= L; ++l)
                for (int m = -l; m <= l; ++m) {
                    if (Qdims == 2 && ((l + m) % 2 + 2) % 2 == 1) continue;
                    B.pl[li] = l; B.pm[li] = m;
                    for (int n = 0; n < Nd; ++n) {
                        B.M[n + (size_t)Nd * li] = (double)(2 * l + 1) / (4 * PI) * R[n][l][m + l];
                        B.D[li + (size_t)B.Np * n] = w[n] * R[n][l][m + l];
                    }
                    ++li;
                }
        } else if (gal) {
            int gt = (type == "galerkin-d") ? 1 : 2;
            int Lg = Lb;
            bool ok = gram_schmidt(Nd, Lg, R, w, Qdims, gt, B);
            while (!ok) {
                ++Lg;
                if (Lg > Lb + Nd + 10) fail("Galerkin: unable to find enough independent spherical harmonics");
                for (int n = 0; n < Nd; ++n) R[n] = real_sh_up_to_L(Lg, mu[n], phi[n]);
                ok = gram_schmidt(Nd, Lg, R, w, Qdims, gt, B);
            }
        } else {
            fail("Unknown method.");
        }
    } else {
        fail("Domain dimension should be either 1, 2 or 3.");
    }
    return B;
}

struct SurfBasis {
    int Np = 0;
    vector<double> M, D;      // M: Nd+ x Np, D: Np x Nd+
    vector<long> n_to_np;     // [Nd] 1-based half-range index or 0
    int Ndp = 0;
};

// particle_transport/angular_polynomial_basis.jl:262-436 (half-range basis of one face).
SurfBasis surface_basis(const vector<double>& mu, const vector<double>& eta, const vector<double>& xi,
                        const vector<double>& w, int L, const string& type, int Qdims, int face) {
    int Nd = (int)mu.size();
    SurfBasis S;
    vector<int> idx;
    vector<double> mup, phip;
    if (Qdims == 1) {
        if (face > 1) fail("Unknown surface for 1D angular domain.");
        for (int n = 0; n < Nd; ++n)
            if (face == 0 ? mu[n] > 0 : mu[n] < 0) { idx.push_back(n); mup.push_back(face == 0 ? mu[n] : -mu[n]); }
    } else {
        for (int n = 0; n < Nd; ++n) {
            double a, ph;
            switch (face / 2) {
                case 0: a = mu[n]; ph = std::atan2(xi[n], eta[n]); break;
                case 1: a = eta[n]; ph = std::atan2(mu[n], xi[n]); break;
                default: a = xi[n]; ph = std::atan2(eta[n], mu[n]); break;
            }
            if (face % 2 == 0 ? a > 0 : a < 0) {
                idx.push_back(n);
                mup.push_back(face % 2 == 0 ? a : -a);
                phip.push_back(ph);
            }
        }
    }
    int Ndp = (int)idx.size();
    S.Ndp = Ndp;
    vector<double> wp(Ndp);
    for (int i = 0; i < Ndp; ++i) wp[i] = w[idx[i]];
    S.n_to_np.assign(Nd, 0);
    for (int i = 0; i < Ndp; ++i) S.n_to_np[idx[i]] = i + 1;
    bool gal = (type == "galerkin-m" || type == "galerkin-d");
    int Lb = gal ? std::max(L, galerkin_minimum_order(Ndp, Qdims)) : L;
    if (Qdims == 1) {
        vector<vector<double>> Pl(Ndp);
        for (int i = 0; i < Ndp; ++i) Pl[i] = half_range_legendre_up_to_L(Lb, mup[i]);
        if (type == "standard") {
            S.Np = L + 1;
            S.M.assign((size_t)Ndp * S.Np, 0); S.D.assign((size_t)S.Np * Ndp, 0);
            for (int l = 0; l <= L; ++l)
                for (int i = 0; i < Ndp; ++i) {
                    S.M[i + (size_t)Ndp * l] = Pl[i][l];
                    S.D[l + (size_t)S.Np * i] = wp[i] * Pl[i][l];
                }
        } else if (type == "galerkin-m") {
            S.Np = Ndp;
            S.M.assign((size_t)Ndp * Ndp, 0);
            for (int l = 0; l < Ndp; ++l)
                for (int i = 0; i < Ndp; ++i) S.M[i + (size_t)Ndp * l] = Pl[i][l];
            S.D = invert(S.M, Ndp);
        } else if (type == "galerkin-d") {
            S.Np = Ndp;
            S.D.assign((size_t)Ndp * Ndp, 0);
            for (int l = 0; l < Ndp; ++l)
                for (int i = 0; i < Ndp; ++i) S.D[l + (size_t)Ndp * i] = wp[i] * Pl[i][l];
            S.M = invert(S.D, Ndp);
        } else {
            fail("Unknown method.");
        }
    } else {
        vector<Harm> R(Ndp);
        for (int i = 0; i < Ndp; ++i) R[i] = real_half_range_sh_up_to_L(Lb, std::fabs(mup[i]), phip[i]);
        if (type == "standard") {
            S.Np = (Qdims == 2) ? (L + 1) * (L + 2) / 2 : (L + 1) * (L + 1);
            S.M.assign((size_t)Ndp * S.Np, 0); S.D.assign((size_t)S.Np * Ndp, 0);
            int li = 0;
            for (int l = 0; l <= L; ++l)
                for (int m = -l; m <= l; ++m) {
                    if (Qdims == 2 && ((l + m) % 2 + 2) % 2 == 1) continue;
                    for (int i = 0; i < Ndp; ++i) {
                        S.M[i + (size_t)Ndp * li] = R[i][l][m + l];
                        S.D[li + (size_t)S.Np * i] = wp[i] * R[i][l][m + l];
                    }
                    ++li;
                }
        } else if (gal) {
            int gt = (type == "galerkin-d") ? 1 : 3;
            int Lg = Lb;
            Basis B;
            bool ok = gram_schmidt(Ndp, Lg, R, wp, Qdims, gt, B);
            while (!ok) {
                ++Lg;
                if (Lg > Lb + Ndp + 10) fail("Surface Galerkin: unable to find enough independent half-range harmonics");
                for (int i = 0; i < Ndp; ++i) R[i] = real_half_range_sh_up_to_L(Lg, std::fabs(mup[i]), phip[i]);
                ok = gram_schmidt(Ndp, Lg, R, wp, Qdims, gt, B);
            }
            S.Np = B.Np; S.M = B.M; S.D = B.D;
        } else {
            fail("Unknown method.");
        }
    }
    return S;
}

// ---------------------------------------------------------------------------------------
// Problem and solver state
// ---------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------
// Angular Fokker-Planck operator in moment space (fokker_planck_scattering_matrix.jl:29-40)
// ---------------------------------------------------------------------------------------

// C = A B for column-major A (r x k) and B (k x c).
vector<double> matmul(const vector<double>& A, const vector<double>& B, int r, int k, int c) {
    vector<double> C((size_t)r * c, 0.0);
    for (int j = 0; j < c; ++j)
        for (int q = 0; q < k; ++q) {
            const double b = B[q + (size_t)k * j];
            for (int i = 0; i < r; ++i) C[i + (size_t)r * j] += A[i + (size_t)r * q] * b;
        }
    return C;
}

// ℳ_p = Dn ℳ_n Mn for a direction-space operator ℳ_n (Nd x Nd).
vector<double> to_moments(const Basis& B, const vector<double>& Mn_dir, int Nd) {
    vector<double> t = matmul(B.D, Mn_dir, B.Np, Nd, Nd);
    return matmul(t, B.M, B.Np, Nd, B.Np);
}

// particle_transport/fokker_planck_galerkin.jl:27-51 -- diag(-l(l+1)) mapped to directions,
// shifted by λ₀ = max(0, max_n -ℳ_n[n,n]) on the diagonal, and mapped back.
vector<double> fp_galerkin(const Basis& B, int Nd, double& lam0) {
    const int P = B.Np;
    vector<double> Mp0((size_t)P * P, 0.0);
    for (int p = 0; p < P; ++p) Mp0[p + (size_t)P * p] = -(double)(B.pl[p] * (B.pl[p] + 1));
    vector<double> Mn_dir = matmul(matmul(B.M, Mp0, Nd, P, P), B.D, Nd, P, Nd);
    lam0 = 0.0;
    for (int n = 0; n < Nd; ++n)
        if (Mn_dir[n + (size_t)Nd * n] < -lam0) lam0 = -Mn_dir[n + (size_t)Nd * n];
    for (int n = 0; n < Nd; ++n) Mn_dir[n + (size_t)Nd * n] += lam0;
    return to_moments(B, Mn_dir, Nd);
}

// particle_transport/fokker_planck_finite_difference.jl, Qdims = 1 branch: Morel's weighted
// three-point scheme on the directions sorted by mu (β_{1/2} = 0, β_{n+1/2} = β_{n-1/2} - 2 w_n mu_n),
// with λ₀ = max_n C_n added on the diagonal.
vector<double> fp_fd_1d(const vector<double>& mu, const vector<double>& w, double& lam0) {
    const int N = (int)mu.size();
    vector<int> idx(N);
    for (int i = 0; i < N; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return mu[a] < mu[b]; });
    vector<double> m(N), ww(N);
    for (int i = 0; i < N; ++i) { m[i] = mu[idx[i]]; ww[i] = w[idx[i]]; }
    vector<double> beta(N + 1, 0.0), Cm(N, 0.0), Cp(N, 0.0), C(N);
    for (int n = 1; n < N; ++n) beta[n] = beta[n - 1] - 2.0 * ww[n - 1] * m[n - 1];
    for (int n = 1; n < N; ++n) Cm[n] = beta[n] / (ww[n] * (m[n] - m[n - 1]));
    for (int n = 0; n < N - 1; ++n) Cp[n] = beta[n + 1] / (ww[n] * (m[n + 1] - m[n]));
    lam0 = -std::numeric_limits<double>::infinity();
    for (int n = 0; n < N; ++n) { C[n] = Cm[n] + Cp[n]; lam0 = std::max(lam0, C[n]); }
    vector<double> Mn_dir((size_t)N * N, 0.0);
    for (int n = 0; n < N; ++n) {
        Mn_dir[idx[n] + (size_t)N * idx[n]] = lam0 - C[n];
        if (n > 0) Mn_dir[idx[n] + (size_t)N * idx[n - 1]] = Cm[n];
        if (n < N - 1) Mn_dir[idx[n] + (size_t)N * idx[n + 1]] = Cp[n];
    }
    return Mn_dir;
}

// particle_transport/fokker_planck_finite_difference.jl, Qdims = 2 and 3 branches for the
// gauss-legendre-chebychev product set: Morel's polar operator on the mu levels and a periodic
// (Qdims = 3) or one-sided (Qdims = 2, upper half of the azimuth) three-point azimuthal operator
// scaled by γ_n; directions are ordered azimuth fastest.  λ₀ = max |diag| is added on the diagonal.
vector<double> fp_fd_glc(int N, int Qdims, double& lam0) {
    vector<double> mun, wn;
    gauss_legendre(N, mun, wn);
    vector<int> idx(N);
    for (int i = 0; i < N; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return mun[a] < mun[b]; });
    vector<double> mu(N), wi(N);
    for (int i = 0; i < N; ++i) { mu[i] = mun[idx[i]]; wi[i] = wn[idx[i]]; }
    const int Na = Qdims == 2 ? N : 2 * N;  // azimuthal points per mu level
    const double wj = PI / N;
    vector<double> om(Na);
    for (int i = 0; i < Na; ++i) om[i] = PI * ((double)(i + 1) - 0.5) / N;
    vector<double> beta(N + 1, 0.0), d12(N + 1, 0.0), gam(N, 0.0);
    for (int n = 1; n < N; ++n) {
        beta[n] = beta[n - 1] - 2.0 * wi[n - 1] * mu[n - 1];
        d12[n] = (std::sqrt(1.0 - mu[n] * mu[n]) - std::sqrt(1.0 - mu[n - 1] * mu[n - 1])) / (mu[n] - mu[n - 1]);
    }
    for (int n = 0; n < N; ++n) {
        const double cn = (beta[n + 1] * d12[n + 1] - beta[n] * d12[n]) / wi[n];
        const double Kn = 2.0 * (1.0 - mu[n] * mu[n]) + cn * std::sqrt(1.0 - mu[n] * mu[n]);
        gam[n] = (PI * PI * Kn) / (2.0 * N * N * (1.0 - std::cos(PI / N)));
    }
    const int Nd = N * Na;
    vector<double> Mn_dir((size_t)Nd * Nd, 0.0);
    auto at = [&](int n, int i, int m, int j) -> double& {
        return Mn_dir[(i + (size_t)Na * n) + (size_t)Nd * (j + (size_t)Na * m)];
    };
    for (int n = 0; n < N; ++n) {
        const double a = 1.0 / (1.0 - mu[n] * mu[n]) * gam[n] / wj;  // azimuthal factor of level n
        const double pm = n > 0 ? 1.0 / wi[n] * beta[n] / (mu[n] - mu[n - 1]) : 0.0;
        const double pp = n < N - 1 ? 1.0 / wi[n] * beta[n + 1] / (mu[n + 1] - mu[n]) : 0.0;
        double polar;
        if (n == 0) polar = 1.0 / wi[n] * beta[n + 1] / (mu[n + 1] - mu[n]);
        else if (n == N - 1) polar = 1.0 / wi[n] * beta[n] / (mu[n] - mu[n - 1]);
        else polar = 1.0 / wi[n] * (beta[n + 1] / (mu[n + 1] - mu[n]) + beta[n] / (mu[n] - mu[n - 1]));
        for (int i = 0; i < Na; ++i) {
            double& dg = at(n, i, n, i);
            dg -= polar;
            if (Qdims == 2) {
                if (i == 0) dg -= a / (om[i + 1] - om[i]);
                else if (i == Na - 1) dg -= a / (om[i] - om[i - 1]);
                else dg -= a / (om[i] - om[i - 1]) + a / (om[i + 1] - om[i]);
                if (i > 0) at(n, i, n, i - 1) += a / (om[i] - om[i - 1]);
                if (i < Na - 1) at(n, i, n, i + 1) += a / (om[i + 1] - om[i]);
            } else {
                const double lo = i == 0 ? om[i] + 2.0 * PI - om[Na - 1] : om[i] - om[i - 1];
                const double hi = i == Na - 1 ? om[0] + 2.0 * PI - om[i] : om[i + 1] - om[i];
                dg -= a / lo + a / hi;
                at(n, i, n, i == 0 ? Na - 1 : i - 1) += a / lo;
                at(n, i, n, i == Na - 1 ? 0 : i + 1) += a / hi;
            }
            if (n > 0) at(n, i, n - 1, i) += pm;
            if (n < N - 1) at(n, i, n + 1, i) += pp;
        }
    }
    lam0 = 0.0;
    for (int d = 0; d < Nd; ++d) lam0 = std::max(lam0, std::fabs(Mn_dir[d + (size_t)Nd * d]));
    for (int d = 0; d < Nd; ++d) Mn_dir[d + (size_t)Nd * d] += lam0;
    return Mn_dir;
}

// particle_transport/fokker_planck_differential_quadrature.jl (Qdims = 1): (1 - mu²) W² - 2 mu W
// with W the Lagrange differentiation matrix on the nodes (diagonal = minus the row sum), shifted
// by λ₀ = max(0, max_n -ℳ[n,n]).
vector<double> fp_dq_1d(const vector<double>& mu, double& lam0) {
    const int N = (int)mu.size();
    vector<double> Pi(N, 1.0), W((size_t)N * N, 0.0);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < N; ++m)
            if (n != m) Pi[n] *= (mu[n] - mu[m]);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < N; ++m)
            if (n != m) W[n + (size_t)N * m] = Pi[n] / ((mu[n] - mu[m]) * Pi[m]);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < N; ++m)
            if (n != m) W[n + (size_t)N * n] -= W[n + (size_t)N * m];
    vector<double> W2 = matmul(W, W, N, N, N);
    vector<double> Mn_dir((size_t)N * N);
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < N; ++m)
            Mn_dir[n + (size_t)N * m] = (1.0 - mu[n] * mu[n]) * W2[n + (size_t)N * m] - 2.0 * mu[n] * W[n + (size_t)N * m];
    lam0 = 0.0;
    for (int n = 0; n < N; ++n)
        if (Mn_dir[n + (size_t)N * n] < -lam0) lam0 = -Mn_dir[n + (size_t)N * n];
    for (int n = 0; n < N; ++n) Mn_dir[n + (size_t)N * n] += lam0;
    return Mn_dir;
}

// ---------------------------------------------------------------------------------------
// Finite-difference Fokker-Planck operator on the Voronoi tessellation of the unit sphere, for the
// sets other than gauss-legendre-chebychev in 2D/3D (tools/voronoi.jl; fokker_planck_finite_difference.jl
// :139-179 for Qdims = 2, :255-275 for Qdims = 3).
// ---------------------------------------------------------------------------------------

// voronoi_sphere (voronoi.jl:13-205): for each node n, the neighbour across each edge of its Voronoi
// cell, in polygon order.  The cell's vertices come from every triplet (n, j, k): Newton on
// |x|^2 = 1 and the bisector planes of (n, j) and (n, k) from x = node n (tolerance 1e-7, 100 steps,
// skipped when the Jacobian is singular or the iteration does not settle); vertices that agree on a
// 1e-7 grid are merged with the union of their node sets; a vertex stays when n is its nearest node
// (within 1e-7); the polygon is ordered from its first vertex, each next one the first unused vertex
// sharing exactly one node other than n with the previous.  The neighbour across the edge between
// consecutive vertices is the first other node whose polygon holds both.
vector<vector<int>> voronoi_neighbours(const vector<double>& X, const vector<double>& Y, const vector<double>& Z) {
    const int Nd = (int)X.size();
    vector<vector<vector<int>>> index(Nd);
    for (int n = 0; n < Nd; ++n) {
        vector<std::array<double, 3>> vtx;
        vector<vector<int>> pix;
        const double x1 = X[n], y1 = Y[n], z1 = Z[n];
        for (int j = 0; j < Nd; ++j) {
            if (j == n) continue;
            const double x2 = X[j], y2 = Y[j], z2 = Z[j];
            const double a1 = x1 - x2, b1 = y1 - y2, c1 = z1 - z2;
            const double d1 = 0.5 * (x1 * x1 - x2 * x2 + y1 * y1 - y2 * y2 + z1 * z1 - z2 * z2);
            for (int k = 0; k < Nd; ++k) {
                if (k == j || k == n) continue;
                const double x3 = X[k], y3 = Y[k], z3 = Z[k];
                const double a2 = x1 - x3, b2 = y1 - y3, c2 = z1 - z3;
                const double d2 = 0.5 * (x1 * x1 - x3 * x3 + y1 * y1 - y3 * y3 + z1 * z1 - z3 * z3);
                auto F = [&](const double* x, double* f) {
                    f[0] = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - 1.0;
                    f[1] = a1 * x[0] + b1 * x[1] + c1 * x[2] - d1;
                    f[2] = a2 * x[0] + b2 * x[1] + c2 * x[2] - d2;
                };
                double xm[3] = {x1, y1, z1}, x[3] = {x1, y1, z1};
                double eps = std::numeric_limits<double>::infinity();
                int N = 1;
                bool isinv = true;
                while (eps > 1e-7 && N < 100) {
                    const double J[3][3] = {{2 * xm[0], 2 * xm[1], 2 * xm[2]}, {a1, b1, c1}, {a2, b2, c2}};
                    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                    if (det == 0.0) { isinv = false; break; }
                    double Ji[3][3];  // adjugate / det
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1_ = (r + 1) % 3, c2_ = (r + 2) % 3;
                            Ji[r][c] = (J[r1][c1_] * J[r2][c2_] - J[r1][c2_] * J[r2][c1_]) / det;
                        }
                    double f[3];
                    F(x, f);
                    for (int r = 0; r < 3; ++r) x[r] = xm[r] - (Ji[r][0] * f[0] + Ji[r][1] * f[1] + Ji[r][2] * f[2]);
                    F(x, f);
                    eps = std::max(std::fabs(f[0]), std::max(std::fabs(f[1]), std::fabs(f[2])));
                    for (int r = 0; r < 3; ++r) xm[r] = x[r];
                    ++N;
                }
                if (N < 100 && isinv) {
                    vtx.push_back({x[0], x[1], x[2]});
                    pix.push_back({n, j, k});
                }
            }
        }
        // merge duplicates on the 1e-7 grid (round half to even, as Julia's round)
        vector<std::array<double, 3>> v2;
        vector<vector<int>> p2;
        std::map<std::array<long long, 3>, int> seen;
        for (size_t i = 0; i < vtx.size(); ++i) {
            const std::array<long long, 3> key = {(long long)std::nearbyint(vtx[i][0] * 1e7),
                                                  (long long)std::nearbyint(vtx[i][1] * 1e7),
                                                  (long long)std::nearbyint(vtx[i][2] * 1e7)};
            auto it = seen.find(key);
            if (it != seen.end()) {
                vector<int>& u = p2[it->second];
                for (int q : pix[i])
                    if (std::find(u.begin(), u.end(), q) == u.end()) u.push_back(q);
            } else {
                v2.push_back(vtx[i]);
                p2.push_back(pix[i]);
                seen[key] = (int)v2.size() - 1;
            }
        }
        // the vertices whose nearest node is n
        vector<vector<int>> p3;
        for (size_t i = 0; i < v2.size(); ++i) {
            double dn = 0.0, dmin = std::numeric_limits<double>::infinity();
            for (int q = 0; q < Nd; ++q) {
                const double dx = X[q] - v2[i][0], dy = Y[q] - v2[i][1], dz = Z[q] - v2[i][2];
                const double dq = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (q == n) dn = dq;
                dmin = std::min(dmin, dq);
            }
            if (std::fabs(dn - dmin) < 1e-7) {
                vector<int> srt = p2[i];
                std::sort(srt.begin(), srt.end());
                p3.push_back(srt);
            }
        }
        // polygon order
        vector<vector<int>> ord;
        if (!p3.empty()) ord.push_back(p3[0]);
        auto others = [&](const vector<int>& v) {
            vector<int> o;
            for (int q : v)
                if (q != n) o.push_back(q);
            return o;
        };
        for (size_t i = 1; i < p3.size(); ++i) {
            if (ord.size() < i) break;
            const vector<int> a = others(ord[i - 1]);
            for (size_t j = 0; j < p3.size(); ++j) {
                const vector<int> b = others(p3[j]);
                int common = 0;
                for (int q : a)
                    if (std::find(b.begin(), b.end(), q) != b.end()) ++common;
                if (common == 1 && std::find(ord.begin(), ord.end(), p3[j]) == ord.end()) {
                    ord.push_back(p3[j]);
                    break;
                }
            }
        }
        index[n] = ord;
    }
    vector<vector<int>> nb(Nd);
    for (int n = 0; n < Nd; ++n) {
        const int M = (int)index[n].size();
        for (int m = 0; m < M; ++m) {
            const vector<int>& i2 = index[n][m == 0 ? M - 1 : m - 1];
            const vector<int>& i3 = index[n][m];
            int found = -1;
            for (int i = 0; i < Nd && found < 0; ++i) {
                if (i == n) continue;
                const auto& L = index[i];
                if (std::find(L.begin(), L.end(), i2) != L.end() && std::find(L.begin(), L.end(), i3) != L.end())
                    found = i;
            }
            if (found < 0) fail("Voronoi cell edge without a neighbour");
            nb[n].push_back(found);
        }
    }
    return nb;
}

// Least-squares solution of G x = q (G: m x n column-major, full column rank) by Householder QR: the
// pinv(G) q of voronoi.jl, whose check ||pinv(G) G - I|| <= 1e-7 (full column rank) is made here on R.
vector<double> least_squares(vector<double> G, vector<double> q, int m, int n) {
    double rmax = 0.0;
    vector<double> diag(n);
    for (int k = 0; k < n; ++k) {
        double nrm = 0.0;
        for (int i = k; i < m; ++i) nrm += G[i + (size_t)m * k] * G[i + (size_t)m * k];
        nrm = std::sqrt(nrm);
        const double a = G[k + (size_t)m * k] > 0 ? -nrm : nrm;
        diag[k] = a;
        rmax = std::max(rmax, std::fabs(a));
        if (nrm == 0.0) continue;
        vector<double> v(m - k);
        for (int i = k; i < m; ++i) v[i - k] = G[i + (size_t)m * k];
        v[0] -= a;
        double vv = 0.0;
        for (double t : v) vv += t * t;
        if (vv == 0.0) continue;
        for (int c = k; c < n; ++c) {
            double s = 0.0;
            for (int i = k; i < m; ++i) s += v[i - k] * G[i + (size_t)m * c];
            s = 2.0 * s / vv;
            for (int i = k; i < m; ++i) G[i + (size_t)m * c] -= s * v[i - k];
        }
        double s = 0.0;
        for (int i = k; i < m; ++i) s += v[i - k] * q[i];
        s = 2.0 * s / vv;
        for (int i = k; i < m; ++i) q[i] -= s * v[i - k];
    }
    for (int k = 0; k < n; ++k)
        if (!(std::fabs(G[k + (size_t)m * k]) > 1e-10 * rmax))
            fail("The equation system to find γ-parameters does not have a solution.");
    vector<double> x(n);
    for (int k = n - 1; k >= 0; --k) {
        double v = q[k];
        for (int c = k + 1; c < n; ++c) v -= G[k + (size_t)m * c] * x[c];
        x[k] = v / G[k + (size_t)m * k];
    }
    return x;
}

// fokker_planck_weights_3D / _2D (voronoi.jl:207-353) and the operator of
// fokker_planck_finite_difference.jl:139-179, 255-275: one unknown γ per Voronoi edge (n, m), fixed by
// sum_m γ_nm (Ω_k[m] - Ω_k[n]) = -2 w_n Ω_k[n] for every node and axis k (the first moments are
// eigenfunctions of eigenvalue -2); ℳ[n][m] = γ_nm / w_n off the diagonal, rows summing to zero, then
// shifted by λ₀ = max |ℳ[n][n]|.  Qdims = 2 takes the edges of the full-sphere set (X3, Y3, Z3) between
// nodes of the upper hemisphere, mapped onto the folded set (mu, eta, xi, w).
vector<double> fp_fd_voronoi(int Qdims, const vector<double>& mu, const vector<double>& eta, const vector<double>& xi,
                             const vector<double>& w, const vector<double>& X3, const vector<double>& Y3,
                             const vector<double>& Z3, double& lam0) {
    const int Nd = (int)w.size();
    const vector<vector<int>> nb3 = voronoi_neighbours(X3, Y3, Z3);
    const int N3 = (int)X3.size();
    vector<int> to2(N3, -1), to3(Nd, -1);
    if (Qdims == 3) {
        if (N3 != Nd) fail("Fokker-Planck operator does not match the quadrature");
        for (int n = 0; n < Nd; ++n) to2[n] = to3[n] = n;
    } else {
        for (int n3 = 0; n3 < N3; ++n3) {
            int best = 0;
            double bv = 0.0;
            for (int j = 0; j < Nd; ++j) {
                const double sz = Z3[n3] > 0 ? 1.0 : (Z3[n3] < 0 ? -1.0 : 0.0);
                const double e = std::pow(std::fabs(X3[n3] - mu[j]), 2) + std::pow(std::fabs(Y3[n3] - eta[j]), 2) +
                                 std::pow(std::fabs(sz * Z3[n3] - xi[j]), 2);
                if (j == 0 || e < bv) { bv = e; best = j; }
            }
            to2[n3] = best;
            if (Z3[n3] >= 0) to3[best] = n3;
        }
        for (int n = 0; n < Nd; ++n)
            if (to3[n] < 0) fail("Fokker-Planck operator does not match the quadrature");
    }
    auto upper = [&](int n3) { return Qdims == 3 || Z3[n3] >= 0; };
    // edges in order of discovery
    vector<std::pair<int, int>> edges;
    std::map<std::pair<int, int>, int> eid;
    for (int n3 = 0; n3 < N3; ++n3) {
        if (!upper(n3)) continue;
        for (int m3 : nb3[n3]) {
            if (!upper(m3)) continue;
            const std::pair<int, int> e = {std::min(to2[n3], to2[m3]), std::max(to2[n3], to2[m3])};
            if (!eid.count(e)) { eid[e] = (int)edges.size(); edges.push_back(e); }
        }
    }
    const int Ne = (int)edges.size(), Neq = 3 * Nd;
    const double* Om[3] = {mu.data(), eta.data(), xi.data()};
    vector<double> G((size_t)Neq * Ne, 0.0), Q(Neq, 0.0);
    for (int n = 0; n < Nd; ++n)
        for (int k = 0; k < 3; ++k) {
            const int eq = 3 * n + k;
            for (int m3 : nb3[to3[n]]) {
                if (!upper(m3)) continue;
                const int m = to2[m3];
                if (m == n) continue;
                const int e = eid.at({std::min(n, m), std::max(n, m)});
                G[eq + (size_t)Neq * e] -= Om[k][n];
                G[eq + (size_t)Neq * e] += Om[k][m];
                Q[eq] = -2.0 * w[n] * Om[k][n];
            }
        }
    const vector<double> gam = least_squares(G, Q, Neq, Ne);
    vector<double> Mn_dir((size_t)Nd * Nd, 0.0);
    for (int n = 0; n < Nd; ++n)
        for (int m3 : nb3[to3[n]]) {
            if (!upper(m3)) continue;
            const int m = to2[m3];
            if (m == n) continue;
            const double g = gam[eid.at({std::min(n, m), std::max(n, m)})];
            Mn_dir[n + (size_t)Nd * m] += g / w[n];
            Mn_dir[n + (size_t)Nd * n] -= g / w[n];
        }
    lam0 = 0.0;
    for (int d = 0; d < Nd; ++d) lam0 = std::max(lam0, std::fabs(Mn_dir[d + (size_t)Nd * d]));
    for (int d = 0; d < Nd; ++d) Mn_dir[d + (size_t)Nd * d] += lam0;
    return Mn_dir;
}

// ---------------------------------------------------------------------------------------
// Lorentz-force operator of an external magnetic field (electromagnetic_scattering_matrix.jl)
// ---------------------------------------------------------------------------------------

// ferrer_associated_legendre.jl: P_l^m(x) without the (-1)^m factor, by the upward recursion from
// P_m^m = (2m-1)!! (1-x^2)^(m/2).
double ferrer(int l, int m, double x) {
    if (l < 0 || m < 0 || m > l || std::fabs(x) > 1) fail("invalid associated Legendre arguments");
    double Pmm = 1.0, Pmm1 = 1.0, Pmm2 = 1.0;
    if (m > 0) {
        double df = 1.0;
        for (int k = 1; k <= 2 * m - 1; k += 2) df *= k;
        Pmm = df * std::pow(std::sqrt(1.0 - x * x), m);
    }
    if (l == m) return Pmm;
    Pmm1 = (2 * m + 1) * x * Pmm;
    if (l == m + 1) return Pmm1;
    for (int li = m + 2; li <= l; ++li) {
        Pmm2 = ((2 * li - 1) * x * Pmm1 - (li + m - 1) * Pmm) / (li - m);
        Pmm = Pmm1;
        Pmm1 = Pmm2;
    }
    return Pmm2;
}

// electromagnetic_scattering_matrix.jl:43-127 (magnetic field B, no electric field): per group
// ℳ_EM[g] = -𝒯_g D M with 𝒯_g = 2/ΔE_g log((sqrt(Eb_g + 2 mc²) + sqrt(Eb_g)) / (sqrt(Eb_g+1 + 2 mc²) +
// sqrt(Eb_g+1))) (mc² = 0.510999, c = 2.99792458) and M[n][p] the derivative of the harmonic p at
// direction n along the Lorentz rotation, charge c (2l+1)/(4π) [(B_z η - B_y ξ) ∂Y/∂μ + (B_y μ η +
// B_z μ ξ - B_x (1-μ²))/(1-μ²) ∂Y/∂φ], with the polar-axis limits of the reference at |μ| = 1.
// Returns [G][Np][Np] (ℳ_EM[g] column-major at g*Np*Np); λ₀_EM is zero.
vector<double> em_matrix(const double* Bf, double charge, const vector<double>& mu, const vector<double>& eta,
                         const vector<double>& xi, const Basis& B, int G, const double* Eb) {
    const int Nd = (int)mu.size(), P = B.Np;
    const double c = 2.99792458, mc2 = 0.510999, PI = 3.14159265358979323846;
    vector<double> T(G);
    for (int g = 0; g < G; ++g) {
        const double dE = Eb[g] - Eb[g + 1];
        T[g] = 2.0 / dE * std::log((std::sqrt(Eb[g] + 2 * mc2) + std::sqrt(Eb[g])) /
                                   (std::sqrt(Eb[g + 1] + 2 * mc2) + std::sqrt(Eb[g + 1])));
    }
    vector<double> M((size_t)Nd * P, 0.0);  // M[n + Nd*p]
    auto sgn = [](double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
    for (int n = 0; n < Nd; ++n) {
        double phi;
        if (xi[n] != 0) {
            if (eta[n] > 0) phi = sgn(xi[n]) * std::atan(std::fabs(xi[n] / eta[n]));
            else if (eta[n] == 0) phi = sgn(xi[n]) * PI / 2;
            else phi = sgn(xi[n]) * (PI - std::atan(std::fabs(xi[n] / eta[n])));
        } else {
            phi = eta[n] >= 0 ? 0.0 : PI;
        }
        const double u = mu[n];
        for (int p = 0; p < P; ++p) {
            const int l = (int)B.pl[p], m = (int)B.pm[p], am = std::abs(m);
            double lf = 0.0;
            for (int k = 1; k <= l - am; ++k) lf += std::log((double)k);
            for (int k = 1; k <= l + am; ++k) lf -= std::log((double)k);
            const double Clm = std::sqrt((2.0 - (m == 0 ? 1.0 : 0.0)) * std::exp(lf));
            const double Tm = m >= 0 ? std::cos(m * phi) : std::sin(am * phi);
            const double nrm = (2 * l + 1) / (4 * PI) * charge * c;
            double v = 0.0;
            // derivative in μ
            if (std::fabs(u) != 1) {
                if (l > 0) {
                    if (l > am) {
                        const double Plm = ferrer(l, am, u), Pl1 = ferrer(l - 1, am, u);
                        v += ((l + am) * Pl1 - l * u * Plm) / (1 - u * u) * Clm * Tm * nrm * (Bf[2] * eta[n] - Bf[1] * xi[n]);
                    } else if (l == m) {
                        const double Pll = ferrer(l, l, u);
                        v += -l * u * Pll / (1 - u * u) * Clm * Tm * nrm * (Bf[2] * eta[n] - Bf[1] * xi[n]);
                    }
                }
            } else if (am == 1) {
                v += -l * (l + 1) / 2.0 * std::pow(u, l) * Clm * Tm * nrm * (Bf[2] * std::cos(phi) - Bf[1] * std::sin(phi));
            }
            // derivative in φ
            const double Plm = ferrer(l, am, u);
            if (std::fabs(u) != 1) {
                const double fac = (Bf[1] * u * eta[n] + Bf[2] * u * xi[n] - Bf[0] * (1 - u * u));
                if (m >= 0) v += Plm * Clm * -m * std::sin(m * phi) * (2 * l + 1) / (4 * PI) * charge * c / (1 - u * u) * fac;
                else v += Plm * Clm * am * std::cos(am * phi) * (2 * l + 1) / (4 * PI) * charge * c / (1 - u * u) * fac;
            } else {
                const double fac = (Bf[1] * std::cos(phi) + Bf[2] * std::sin(phi));
                if (m == 1) v += l * (l + 1) / 2.0 * std::pow(u, l + 1) * Clm * -m * std::sin(m * phi) * (2 * l + 1) / (4 * PI) * charge * c * fac;
                else if (m == -1) v += l * (l + 1) / 2.0 * std::pow(u, l + 1) * Clm * am * std::cos(am * phi) * (2 * l + 1) / (4 * PI) * charge * c * fac;
            }
            M[n + (size_t)Nd * p] = v;
        }
    }
    // G = -(D M) [P x P], ℳ_EM[g] = 𝒯_g G
    vector<double> out((size_t)G * P * P);
    for (int p = 0; p < P; ++p)
        for (int q = 0; q < P; ++q) {
            double s2 = 0.0;
            for (int n = 0; n < Nd; ++n) s2 += B.D[p + (size_t)P * n] * M[n + (size_t)Nd * q];
            for (int g = 0; g < G; ++g) out[(size_t)g * P * P + p + (size_t)P * q] = T[g] * -s2;
        }
    return out;
}

// fokker_planck_scattering_matrix.jl:29-40 -- ℳ (Np x Np column-major) and λ₀ for one method.
vector<double> fokker_planck_matrix(const string& method, int N, const string& qt, int Qdims,
                                    const vector<double>& mu, const vector<double>& eta, const vector<double>& xi,
                                    const vector<double>& w, const Basis& B, const vector<double>* sphere3,
                                    double& lam0) {
    const int Nd = (int)w.size();
    if (method == "galerkin") return fp_galerkin(B, Nd, lam0);
    vector<double> Mn_dir;
    if (method == "finite-difference") {
        if (Qdims == 1) Mn_dir = fp_fd_1d(mu, w, lam0);
        else if (qt == "gauss-legendre-chebychev") Mn_dir = fp_fd_glc(N, Qdims, lam0);
        else Mn_dir = fp_fd_voronoi(Qdims, mu, eta, xi, w, sphere3[0], sphere3[1], sphere3[2], lam0);
        if ((int)(Mn_dir.size()) != Nd * Nd) fail("Fokker-Planck operator does not match the quadrature");
    } else if (method == "differential-quadrature") {
        if (Qdims != 1) fail("Differential quadrature unavailable in multidimensional geometry.");
        Mn_dir = fp_dq_1d(mu, lam0);
    } else {
        fail("Unknown method to treat the Fokker-Planck term.");
    }
    return to_moments(B, Mn_dir, Nd);
}

// ---------------------------------------------------------------------------------------
// Spatial closure of the 1D cell (scheme_weights.jl:25-90, sn_1D_bte.jl:27-64)
// ---------------------------------------------------------------------------------------

// scheme_weights.jl, Ndims = 1 without CSD: ω[Mx+2] for order K = Mx+1.
vector<double> scheme_weights_1d(const string& scheme, int K) {
    const int Mx = K - 1;
    vector<double> w(Mx + 2, 0.0);
    auto pw = [](int e) { return (e % 2 == 0) ? 1.0 : -1.0; };  // (-1)^e, e >= 0
    if (scheme == "DD" || (scheme == "AWD" && Mx == 0)) {
        w[0] = pw(Mx + 1);
        for (int n = 0; n <= Mx; ++n) w[n + 1] = 1.0 - pw(Mx + 1 - n);
    } else if (scheme == "DG" || (scheme == "AWD" && Mx == 1)) {
        w[0] = 0.0;
        for (int n = 0; n <= Mx; ++n) w[n + 1] = 1.0;
    } else if (scheme == "DG-" || scheme == "DG+") {
        if (Mx < 1) fail(scheme + " scheme require at least a 2nd-order p expansion.");
        w[0] = 0.0;
        for (int n = 0; n < Mx; ++n) w[n + 1] = 1.0;
        w[Mx + 1] = scheme == "DG-" ? (double)Mx / (2.0 * Mx + 1.0) : 1001.0;
    } else {
        fail("Closure relation not implemented yet.");
    }
    return w;
}

// Dense solve S x = b (K x K, row-major S), LU with partial pivoting: the role of Julia's `\`.
vector<double> lu_solve(vector<double> S, vector<double> b, int K) {
    for (int k = 0; k < K; ++k) {
        int piv = k;
        for (int i = k + 1; i < K; ++i)
            if (std::fabs(S[i * K + k]) > std::fabs(S[piv * K + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < K; ++j) std::swap(S[k * K + j], S[piv * K + j]);
            std::swap(b[k], b[piv]);
        }
        if (S[k * K + k] == 0.0) fail("singular cell system");
        for (int i = k + 1; i < K; ++i) {
            const double f = S[i * K + k] / S[k * K + k];
            for (int j = k; j < K; ++j) S[i * K + j] -= f * S[k * K + j];
            b[i] -= f * b[k];
        }
    }
    vector<double> x(K);
    for (int i = K - 1; i >= 0; --i) {
        double v = b[i];
        for (int j = i + 1; j < K; ++j) v -= S[i * K + j] * x[j];
        x[i] = v / S[i * K + i];
    }
    return x;
}

// The adaptive weighted diamond closure (scheme "AWD", adaptive.jl): a cell's weights follow from
// its incoming face fluxes and source moments, cell by cell in every sweep.
// AWD0, order 1 on n axes (adaptive.jl:36-54, 127-158, 257-297): the diamond estimate
//   phi_DD = (Q_1 + 2 h_1 f_1 + ... + 2 h_n f_n) / (2 h_1 + ... + 2 h_n + St),
//   u_a = (f_a - phi_DD) / phi_DD,  P_a = 1 when 3|u_a| <= 1, else 1/(3|u_a|); 0 when negative or NaN;
//   w_a = [-P_a, 1 + P_a].
void awd0_weights(int n, const double* h, const double* f, double Q1, double St, double* const* w) {
    double num = Q1, den = 2.0 * h[0];
    for (int a = 0; a < n; ++a) num += 2.0 * h[a] * f[a];
    for (int a = 1; a < n; ++a) den += 2.0 * h[a];
    const double phi_dd = num / (den + St);
    for (int a = 0; a < n; ++a) {
        const double u = (f[a] - phi_dd) / phi_dd;
        double P = 3.0 * std::fabs(u) <= 1.0 ? 1.0 : 1.0 / (3.0 * std::fabs(u));
        if (P < 0.0 || std::isnan(P)) P = 0.0;
        w[a][0] = -P;
        w[a][1] = 1.0 + P;
    }
}
// AWD1, order 2 in 1D (adaptive.jl:56-73): dQ = sqrt3 sx (Q_2 - sqrt3 hx sx f) / (Q_1 + hx f),
//   Qx = (3 + dQ)/(3 (3 - dQ)) when dQ > 3/2, -1/dQ when dQ < -1, else 1;  w = [0, 1, Qx].
void awd1_weights_1d(double hx, double sx, double f, const double* Qn, double* w) {
    const double r3 = std::sqrt(3.0);
    const double dQ = r3 * sx * (Qn[1] - r3 * hx * sx * f) / (Qn[0] + hx * f);
    double Qx;
    if (dQ > 1.5) Qx = (3.0 + dQ) / (3.0 * (3.0 - dQ));
    else if (dQ < -1.0) Qx = -1.0 / dQ;
    else Qx = 1.0;
    w[0] = 0.0;
    w[1] = 1.0;
    w[2] = Qx + 0.0;
}

// constant_linear.jl: the weights (Qx, Qy) and (Tx, Ty) of the coupled linear closure from the centroid
// (x, y) of the cell's constant part in [-1/2, 1/2]^2, by subdomain (A, B, C types and their signs).
namespace cl {
double y_in(double x) {
    const double r = std::sqrt(36 * x * x - 60 * x + 9);
    return (-6 * x - 3 + r) / (36 * x - 30 - 6 * r);
}
double y_out(double x) { return (6 * x * x - 3 * x) / (6 * x + 1); }
// subdomains: 0 A++ 1 A+- 2 A-+ 3 A-- 4 Bx+ 5 Bx- 6 By+ 7 By- 8 C++ 9 C+- 10 C-+ 11 C--
int which(double x, double y) {
    if (x < -0.5 || x > 0.5 || y < -0.5 || y > 0.5)
        fail("The centroid position of the finite element constant part is out of the finite element domain.");
    const double s6 = 1.0 / 6.0;
    if (-0.5 <= x && x <= -s6) {
        if (-0.5 <= y && y <= -s6) return 11;
        if (-s6 < y && y < s6) return 5;
        return 10;
    }
    if (-s6 < x && x <= 0) {
        const double yo = y_out(-x), yi = y_in(-x);
        if (-0.5 <= y && y <= yo) return 7;
        if (yo < y && y < -yi) return 3;
        if (-yi <= y && y <= yi) return 5;
        if (yi < y && y < -yo) return 2;
        return 6;
    }
    if (0 < x && x < s6) {
        const double yo = y_out(x), yi = y_in(x);
        if (-0.5 <= y && y <= yo) return 7;
        if (yo < y && y < -yi) return 1;
        if (-yi <= y && y <= yi) return 4;
        if (yi < y && y < -yo) return 0;
        return 6;
    }
    if (-0.5 <= y && y <= -s6) return 9;
    if (-s6 < y && y < s6) return 4;
    return 8;
}
void dA(double x, double y, double& d1, double& d2) {
    if (std::fabs(x) < 1e-10 || std::fabs(y) < 1e-10) { d1 = 0; d2 = 0; return; }
    using cd = std::complex<double>;
    const double a = 2 * y / x, b = -6 * y / x + 3 - 6 * y, c = 9 * y / (2 * x) - 4.5 - 9 * x + 9 * y, d = 12 * x;
    const double D0 = b * b - 3 * a * c, D1 = 2 * b * b * b - 9 * a * b * c + 27 * a * a * d;
    const cd xi((-1.0) / 2, -std::sqrt(3.0) / 2);
    const cd C = std::pow((D1 + std::sqrt(cd(D1 * D1 - 4 * D0 * D0 * D0, 0.0))) / 2.0, 1.0 / 3.0) * xi;
    if (D0 == 0 && D1 == 0) d1 = -b / (3 * a);
    else d1 = std::real(-(b + C + D0 / C) / (3 * a));
    d2 = (2 * d1 - 3) * y / (2 * x) + 1.5;
}
void dB(double x, double y, double& d1, double& d2) {
    d1 = (12 * y * y + (6 - 12 * x) * y + 2 * x) / (12 * y * y + 1);
    d2 = (12 * y * y - (6 - 12 * x) * y + 2 * x) / (12 * y * y + 1);
}
void dC(double x, double y, double& d1, double& d2) { d1 = 3 * x - 0.5; d2 = 3 * y - 0.5; }
double QAp(double i, double j) { return (6 - 2 * i * i * j) / (3 * i * i * i * j + 6 * (1 - i * i * j)); }
double QAm(double i, double j) { return (2 * i * i * j + 6 * (1 - i * j)) / (3 * i * i * i * j + 6 * (1 - i * i * j)); }
double QBp(double i, double j) {
    if (std::fabs(i - j) < 1e-5) return (1 + i) / ((1 - i) * (1 + 2 * i));
    return (1 - (i * i + j * j + i * j) / 3) / (1 - (i * i + j * j + i * j) + 0.5 * (i * i * i + j * j * j + i * i * j + i * j * j));
}
double QB0(double i, double j) {
    if (std::fabs(i - j) < 1e-5 && std::fabs(i - 1) < 1e-5) return 1.0;
    return 2.0 / 3.0 * (i + 2 * j - 3) / (i + j - 2);
}
double QBm(double i, double j) {
    if (std::fabs(i - j) < 1e-5) return 1 / (1 + 2 * i);
    return (1 - i - j + (i * i + j * j + i * j) / 3) / (1 - (i * i + j * j + i * j) + 0.5 * (i * i * i + j * j * j + i * i * j + i * j * j));
}
double QCp(double i) { return 2 * (2 + i) / (3 * (1 - i * i)); }
double QCm(double i) { return 2 / (3 * (1 + i)); }
double TAp(double i, double j) {
    const double den = i * i * i * j - 2 * i * i * j + 2;
    return -i * i * j * (i * i * i * j * j + 12 * i * j - 30 * i - 20 * j + 40) / (5 * den * den);
}
double TAm(double i, double j) {
    const double den = i * i * i * j - 2 * i * i * j + 2;
    return -i * j * (i * i * i * i * j * j - 8 * i * i * i * j * j + (10 * j * j - 12 * j + 30) * i * i + (40 * j - 80) * i - 40 * j + 60) /
           (5 * den * den);
}
double TBden(double i, double j) {
    const double v = i * i * i + (j - 2) * i * i + (j * j - 2 * j) * i + j * j * j - 2 * j * j + 2;
    return 5 * v * v;
}
double TBp(double i, double j) {
    if (std::fabs(i + j - 2) < 1e-5) return 0.0;
    return -(i - j) * ((i * i * i * i + 4 * i * i * i * j + 10 * i * i * j * j + 4 * i * j * j * j + j * j * j * j) -
                       6 * (3 * i * i + 4 * i * j + 3 * j * j) + 20 * (i + j)) / TBden(i, j);
}
double TBm(double i, double j) {
    if (std::fabs(i - j) < 1e-5) return 0.0;
    return (i - j) * ((i * i * i * i + 4 * i * i * i * j + 10 * i * i * j * j + 4 * i * j * j * j + j * j * j * j) -
                      8 * (i * i * i + 4 * i * i * j + 4 * i * j * j + j * j * j) + 4 * (7 * i * i + 16 * i * j + 7 * j * j) -
                      40 * (i + j) + 20) / TBden(i, j);
}
double TB0(double i, double j) {
    if (std::fabs(i - j) < 1e-5) return 0.0;
    return -(i - j) * (i * i + 6 * i * j + 3 * j * j - 8 * i - 12 * j + 10) / (5 * (i + j - 2) * (i + j - 2));
}
double TCp(double i, double j) {
    if (std::fabs(i - j) < 1e-5) return -(i * i + 6 * i + 3) / (5 * (1 + i) * (1 + i));
    if (std::fabs(i - 1) < 1e-5) return -std::numeric_limits<double>::infinity();
    return -(i * i + 6 * i + 3) * (j - 1) / (5 * (1 + i) * (1 + i) * (i - 1));
}
double TCm(double i, double j) { return -(i - 1) * (j - 1) / (5 * (1 + i) * (1 + i)); }

// constant_linear.jl:19-138: subdomain, δ lengths (clamped to [0, 1] within 1e-7), then Q and T
void constant_linear(double x, double y, double* Q, double* T) {
    const int sub = which(x, y);
    double d1 = 0, d2 = 0;
    switch (sub) {
        case 0: dA(x, y, d1, d2); break;
        case 1: dA(x, -y, d1, d2); break;
        case 2: dA(-x, y, d1, d2); break;
        case 3: dA(-x, -y, d1, d2); break;
        case 4: dB(x, y, d1, d2); break;
        case 5: dB(-x, y, d1, d2); break;
        case 6: dB(y, x, d1, d2); break;
        case 7: dB(-y, x, d1, d2); break;
        case 8: dC(x, y, d1, d2); break;
        case 9: dC(x, -y, d1, d2); break;
        case 10: dC(-x, y, d1, d2); break;
        default: dC(-x, -y, d1, d2); break;
    }
    auto clamp01 = [](double& d) {
        if (d > 1 && d < 1.0000001) d = 1;
        if (d < 0 && d > -0.0000001) d = 0;
    };
    clamp01(d1);
    clamp01(d2);
    if (!std::isfinite(d1) || !std::isfinite(d2) || d1 < 0 || d1 > 1 || d2 < 0 || d2 > 1)
        fail("constant_linear: δ lengths out of [0, 1]");
    Q[0] = Q[1] = T[0] = T[1] = 0.0;
    switch (sub) {
        case 0: Q[0] += QAp(d1, d2); Q[1] += QAp(d2, d1); T[0] += TAp(d1, d2); T[1] += TAp(d2, d1); break;
        case 1: Q[0] += QAp(d1, d2); Q[1] += QAm(d2, d1); T[0] -= TAp(d1, d2); T[1] -= TAm(d2, d1); break;
        case 2: Q[0] += QAm(d1, d2); Q[1] += QAp(d2, d1); T[0] -= TAm(d1, d2); T[1] -= TAp(d2, d1); break;
        case 3: Q[0] += QAm(d1, d2); Q[1] += QAm(d2, d1); T[0] += TAm(d1, d2); T[1] += TAm(d2, d1); break;
        case 4: Q[0] += QBp(d1, d2); Q[1] += QB0(d1, d2); T[0] += TBp(d1, d2); T[1] += TB0(d1, d2); break;
        case 5: Q[0] += QBm(d1, d2); Q[1] += QB0(d1, d2); T[0] += TBm(d1, d2); T[1] -= TB0(d1, d2); break;
        case 6: Q[0] += QB0(d1, d2); Q[1] += QBp(d1, d2); T[0] += TB0(d1, d2); T[1] += TBp(d1, d2); break;
        case 7: Q[0] += QB0(d1, d2); Q[1] += QBm(d1, d2); T[0] -= TB0(d1, d2); T[1] += TBm(d1, d2); break;
        case 8: Q[0] += QCp(d1); Q[1] += QCp(d2); T[0] += TCp(d1, d2); T[1] += TCp(d2, d1); break;
        case 9: Q[0] += QCp(d1); Q[1] += QCm(d2); T[0] -= TCp(d1, d2); T[1] -= TCm(d2, d1); break;
        case 10: Q[0] += QCm(d1); Q[1] += QCp(d2); T[0] -= TCm(d1, d2); T[1] -= TCp(d2, d1); break;
        default: Q[0] += QCm(d1); Q[1] += QCm(d2); T[0] += TCm(d1, d2); T[1] += TCm(d2, d1); break;
    }
    for (int k = 0; k < 2; ++k)
        if (!std::isfinite(Q[k]) || !std::isfinite(T[k])) fail("constant_linear: weights not finite");
}
}  // namespace cl

// The coupled linear AWD1 of a 2D cell of orders (2, 2) with full coupling (adaptive.jl:161-232):
// per axis the slope of the cell's source and inflows estimates the centroid of its constant part,
// which sets (Q, T) = constant_linear(x_CM, y_CM) and the weights
//   wx[1][m][m] = 1, wx[2][m][m] = Qx, wx[2][0][1] = sy Tx / sqrt3 (likewise wy with Qy, Ty, sx),
// the other entries those of DG (0 for the inflow weight wx[0][m][m]).
void awd1_weights_2d(double hx, double hy, double sx, double sy, const double* fx, const double* fy,
                     const double* Qn, double wx[3][2][2], double wy[3][2][2]) {
    const double r3 = std::sqrt(3.0);
    double xcm[2] = {0.0, 0.0};
    for (int i = 0; i < 2; ++i) {
        double p0, p1, dQ;
        if (i == 0) {
            p0 = Qn[0] + hx * fx[0] + hy * fy[0];
            p1 = Qn[1] - r3 * hx * sx * fx[0] + hy * fy[1];
            dQ = sx * r3 * p1;
        } else {
            p0 = Qn[0] + hx * fx[0] + hy * fy[0];
            p1 = Qn[2] + hx * fx[1] - r3 * hy * sy * fy[0];
            dQ = sy * r3 * p1;
        }
        if (p0 != 0) dQ = dQ / p0;
        else if (std::fabs(dQ) < 1e-8) dQ = 0;
        else dQ = std::numeric_limits<double>::infinity() * (dQ > 0 ? 1.0 : -1.0);
        if (std::isnan(dQ)) dQ = 0;
        if (dQ > 2.999) dQ = 2.999;
        if (dQ < -2.999) dQ = -2.999;
        const double xmax = 0.49, sg = dQ > 0 ? 1.0 : (dQ < 0 ? -1.0 : 0.0), dQt = dQ;
        dQ += 1.5 * sg;
        if (dQ > 1.5) {
            if (dQ > 3) xcm[i] = xmax;
            else xcm[i] = std::min(xmax, (2 * dQ - 3 + std::sqrt(12 * dQ * dQ - 27)) / (4 * (3 * dQ)));
        }
        dQ = dQt + sg;
        if (dQ < -1) xcm[i] = std::max(-xmax, (1 + dQ) / 4);
        if (xcm[i] < -xmax) xcm[i] = -xmax;
        if (xcm[i] > xmax) xcm[i] = xmax;
    }
    double Q[2], T[2];
    cl::constant_linear(xcm[0], xcm[1], Q, T);
    for (int n = 0; n < 3; ++n)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) wx[n][a][b] = wy[n][a][b] = (a == b && n > 0) ? 1.0 : 0.0;
    wx[1][0][0] = 1; wx[2][0][0] = Q[0]; wx[1][1][1] = 1; wx[2][1][1] = Q[0]; wx[2][0][1] = sy * T[0] / r3;
    wy[1][0][0] = 1; wy[2][0][0] = Q[1]; wy[1][1][1] = 1; wy[2][1][1] = Q[1]; wy[2][0][1] = sx * T[1] / r3;
}

// sn_2D_bte.jl:35-116 for orders (2, 2) with full coupling and weight tensors w[n][m][k] (the
// reference's ω[n+1, m+1, k+1]): the streaming terms couple through w[jx+1][jy][iy], the inflow uses
// w[0][m][m], and the outgoing face moment m gathers sum_k w[jx+1][m][k].
void flux_2d_bte_tensor(double mu, double eta, double St, double dx, double dy, const double* Qn, double* fx,
                        double* fy, double wx[3][2][2], double wy[3][2][2], double* psi) {
    const double sx = mu > 0 ? 1.0 : (mu < 0 ? -1.0 : 0.0), sy = eta > 0 ? 1.0 : (eta < 0 ? -1.0 : 0.0);
    const double hx = std::fabs(mu) / dx, hy = std::fabs(eta) / dy;
    auto C = [](int i) { return std::sqrt(2.0 * i - 1.0); };  // 1-based
    auto spow = [](double s, int e) { double v = 1.0; for (int k = 0; k < e; ++k) v *= s; return v; };
    const int O = 2, Nm = 4;
    vector<double> S((size_t)Nm * Nm, 0.0), Q(Nm, 0.0);
    for (int ix = 1; ix <= O; ++ix)
        for (int jx = 1; jx <= O; ++jx)
            for (int iy = 1; iy <= O; ++iy)
                for (int jy = 1; jy <= O; ++jy) {
                    const int i = O * (iy - 1) + ix - 1, j = O * (jy - 1) + jx - 1;
                    double& e = S[(size_t)i * Nm + j];
                    if (i == j) e += St;
                    if (iy == jy && ix >= jx + 1) e -= C(ix) * hx * sx * C(jx) * (1.0 - spow(-1.0, ix - jx));
                    e += C(ix) * hx * spow(sx, ix - 1) * C(jx) * spow(sx, jx - 1) * wx[jx][jy - 1][iy - 1];
                    if (ix == jx && iy >= jy + 1) e -= C(iy) * hy * sy * C(jy) * (1.0 - spow(-1.0, iy - jy));
                    e += C(iy) * hy * spow(sy, iy - 1) * C(jy) * spow(sy, jy - 1) * wy[jy][jx - 1][ix - 1];
                }
    for (int jx = 1; jx <= O; ++jx)
        for (int jy = 1; jy <= O; ++jy) {
            const int j = O * (jy - 1) + jx - 1;
            Q[j] += Qn[j];
            Q[j] -= C(jx) * hx * (spow(sx, jx - 1) * wx[0][jy - 1][jy - 1] - spow(-sx, jx - 1)) * fx[jy - 1];
            Q[j] -= C(jy) * hy * (spow(sy, jy - 1) * wy[0][jx - 1][jx - 1] - spow(-sy, jy - 1)) * fy[jx - 1];
        }
    vector<double> x = lu_solve(S, Q, Nm);
    for (int j = 0; j < Nm; ++j) psi[j] = x[j];
    for (int jx = 1; jx <= O; ++jx)
        for (int jy = 1; jy <= O; ++jy) {
            const int j = O * (jy - 1) + jx - 1;
            if (jx == 1) fx[jy - 1] = wx[0][jy - 1][jy - 1] * fx[jy - 1];
            if (jy == 1) fy[jx - 1] = wy[0][jx - 1][jx - 1] * fy[jx - 1];
            for (int iy = 1; iy <= O; ++iy) fx[jy - 1] += C(jx) * spow(sx, jx - 1) * wx[jx][jy - 1][iy - 1] * x[j];
            for (int ix = 1; ix <= O; ++ix) fy[jx - 1] += C(jy) * spow(sy, jy - 1) * wy[jy][jx - 1][ix - 1] * x[j];
        }
}

// sn_1D_bte.jl:27-64: the K spatial Legendre moments of the cell's angular flux and the outgoing
// face flux.  C[i] = sqrt(2i-1) (scheme_weights.jl:409-410).  adapt: the AWD weights of the cell
// replace om (not for a tangential direction, sn_1D_bte.jl:39).
void flux_1d_bte(double mu, double St, double dx, const double* Qn, double& fx, int K, const double* om_in,
                 double* psi, bool adapt = false) {
    double sx = mu > 0 ? 1.0 : (mu < 0 ? -1.0 : 0.0);
    const bool tang = std::fabs(mu) <= 1e-10;
    if (tang) { sx = 1.0; mu = 0.0; }
    const double hx = std::fabs(mu) / dx;
    double oma[3];
    const double* om = om_in;
    if (adapt && !tang) {
        if (K == 1) {
            double* wp[1] = {oma};
            awd0_weights(1, &hx, &fx, Qn[0], St, wp);
        } else if (K == 2) {
            awd1_weights_1d(hx, sx, fx, Qn, oma);
        } else {
            fail("Adaptive scheme is not available for Ox=" + std::to_string(K) + " order flux expansion over the finite-element.");
        }
        om = oma;
    }
    auto C = [](int i) { return std::sqrt(2.0 * i - 1.0); };  // 1-based
    auto spow = [](double s, int e) { double v = 1.0; for (int k = 0; k < e; ++k) v *= s; return v; };
    vector<double> S((size_t)K * K, 0.0), Q(K, 0.0);
    for (int ix = 1; ix <= K; ++ix)
        for (int jx = 1; jx <= K; ++jx) {
            double& e = S[(ix - 1) * K + (jx - 1)];
            if (ix == jx) e += St;
            if (ix >= jx + 1) e -= C(ix) * hx * sx * C(jx) * (1.0 - spow(-1.0, ix - jx));
            e += C(ix) * hx * spow(sx, ix - 1) * C(jx) * spow(sx, jx - 1) * om[jx];
        }
    for (int jx = 1; jx <= K; ++jx) {
        Q[jx - 1] += Qn[jx - 1];
        Q[jx - 1] -= C(jx) * hx * (spow(sx, jx - 1) * om[0] - spow(-sx, jx - 1)) * fx;
    }
    vector<double> x = lu_solve(S, Q, K);
    for (int j = 0; j < K; ++j) psi[j] = x[j];
    if (!tang) {
        double f = om[0] * fx;
        for (int jx = 1; jx <= K; ++jx) f += C(jx) * spow(sx, jx - 1) * om[jx] * x[jx - 1];
        fx = f;
    }
}

// sn_2D_bte.jl:27-116 (non-adaptive): the Nm spatial moments of the cell (Ox*Oy with full coupling,
// Ox + Oy - 1 without) and the outgoing face moments fx[Oy], fy[Ox] (in: incoming).  The closure
// weights of each axis are those of scheme_weights_1d (scheme_weights.jl fills the diagonal only).
struct Cell2 {
    int Ox, Oy, Nm;
    bool fc;
    // basis index of (ix, iy) (0-based), or -1 when the partial coupling drops it
    int idx(int ix, int iy) const {
        if (fc) return Ox * iy + ix;
        if (ix > 0 && iy > 0) return -1;
        return iy > 0 ? Ox - 1 + iy : ix;
    }
};
void flux_2d_bte(double mu, double eta, double St, double dx, double dy, const double* Qn, double* fx, double* fy,
                 const Cell2& cl, const double* wx_in, const double* wy_in, double* psi, bool adapt = false) {
    const double sx = mu > 0 ? 1.0 : (mu < 0 ? -1.0 : 0.0), sy = eta > 0 ? 1.0 : (eta < 0 ? -1.0 : 0.0);
    const double hx = std::fabs(mu) / dx, hy = std::fabs(eta) / dy;
    const int Ox = cl.Ox, Oy = cl.Oy, Nm = cl.Nm;
    double wxa[2], wya[2];
    const double *wx = wx_in, *wy = wy_in;
    if (adapt && Ox == 2 && Oy == 2 && cl.fc) {  // adaptive.jl:161-232: the coupled linear AWD1
        double wxt[3][2][2], wyt[3][2][2];
        awd1_weights_2d(hx, hy, sx, sy, fx, fy, Qn, wxt, wyt);
        flux_2d_bte_tensor(mu, eta, St, dx, dy, Qn, fx, fy, wxt, wyt, psi);
        return;
    }
    if (adapt) {  // adaptive.jl:121-160: AWD0 for orders (1, 1)
        if (Ox != 1 || Oy != 1)
            fail("the adaptive closure in 2D: orders (1, 1), or (2, 2) with full coupling");
        const double h[2] = {hx, hy}, f[2] = {fx[0], fy[0]};
        double* wp[2] = {wxa, wya};
        awd0_weights(2, h, f, Qn[0], St, wp);
        wx = wxa;
        wy = wya;
    }
    auto C = [](int i) { return std::sqrt(2.0 * i - 1.0); };  // 1-based
    auto spow = [](double s, int e) { double v = 1.0; for (int k = 0; k < e; ++k) v *= s; return v; };
    vector<double> S((size_t)Nm * Nm, 0.0), Q(Nm, 0.0);
    for (int ix = 1; ix <= Ox; ++ix)
        for (int jx = 1; jx <= Ox; ++jx)
            for (int iy = 1; iy <= Oy; ++iy)
                for (int jy = 1; jy <= Oy; ++jy) {
                    const int i = cl.idx(ix - 1, iy - 1), j = cl.idx(jx - 1, jy - 1);
                    if (i < 0 || j < 0) continue;
                    double& e = S[(size_t)i * Nm + j];
                    if (i == j) e += St;
                    if (iy == jy && ix >= jx + 1) e -= C(ix) * hx * sx * C(jx) * (1.0 - spow(-1.0, ix - jx));
                    if (jy == iy) e += C(ix) * hx * spow(sx, ix - 1) * C(jx) * spow(sx, jx - 1) * wx[jx];
                    if (ix == jx && iy >= jy + 1) e -= C(iy) * hy * sy * C(jy) * (1.0 - spow(-1.0, iy - jy));
                    if (jx == ix) e += C(iy) * hy * spow(sy, iy - 1) * C(jy) * spow(sy, jy - 1) * wy[jy];
                }
    for (int jx = 1; jx <= Ox; ++jx)
        for (int jy = 1; jy <= Oy; ++jy) {
            const int j = cl.idx(jx - 1, jy - 1);
            if (j < 0) continue;
            Q[j] += Qn[j];
            Q[j] -= C(jx) * hx * (spow(sx, jx - 1) * wx[0] - spow(-sx, jx - 1)) * fx[jy - 1];
            Q[j] -= C(jy) * hy * (spow(sy, jy - 1) * wy[0] - spow(-sy, jy - 1)) * fy[jx - 1];
        }
    vector<double> x = lu_solve(S, Q, Nm);
    for (int j = 0; j < Nm; ++j) psi[j] = x[j];
    for (int jx = 1; jx <= Ox; ++jx)
        for (int jy = 1; jy <= Oy; ++jy) {
            const int j = cl.idx(jx - 1, jy - 1);
            if (j < 0) continue;
            if (jx == 1) fx[jy - 1] = wx[0] * fx[jy - 1];
            if (jy == 1) fy[jx - 1] = wy[0] * fy[jx - 1];
            fx[jy - 1] += C(jx) * spow(sx, jx - 1) * wx[jx] * x[j];
            fy[jx - 1] += C(jy) * spow(sy, jy - 1) * wy[jy] * x[j];
        }
}

// sn_3D_bte.jl:34-152 (non-adaptive): Nm = Ox*Oy*Oz moments with full coupling, Ox + Oy + Oz - 2
// without; the x, y and z faces carry the moments over (y, z), (x, z) and (x, y).
struct Cell3 {
    int O[3];
    bool fc;
    int Nm, F[3];  // cell moments, face moments per axis
    Cell3(int ox, int oy, int oz, bool full) : O{ox, oy, oz}, fc(full) {
        Nm = fc ? ox * oy * oz : ox + oy + oz - 2;
        F[0] = fc ? oy * oz : oy + oz - 1;
        F[1] = fc ? ox * oz : ox + oz - 1;
        F[2] = fc ? ox * oy : ox + oy - 1;
    }
    bool valid(int a, int b, int c) const { return fc || (a > 0) + (b > 0) + (c > 0) <= 1; }
    int idx(int a, int b, int c) const {
        if (fc) return O[0] * O[1] * c + O[0] * b + a;
        return a + b + c + (b > 0 ? O[0] - 1 : 0) + (c > 0 ? O[0] - 1 + O[1] - 1 : 0);
    }
    int fx(int b, int c) const { return fc ? O[1] * c + b : b + c + (c > 0 ? O[1] - 1 : 0); }
    int fy(int a, int c) const { return fc ? O[0] * c + a : a + c + (c > 0 ? O[0] - 1 : 0); }
    int fz(int a, int b) const { return fc ? O[0] * b + a : a + b + (b > 0 ? O[0] - 1 : 0); }
};
void flux_3d_bte(const double* om, const double* h, double St, const double* Qn, double* fx, double* fy, double* fz,
                 const Cell3& cl, const double* const* w_in, double* psi, bool adapt = false) {
    double s[3], hh[3];
    for (int a = 0; a < 3; ++a) {
        s[a] = om[a] > 0 ? 1.0 : (om[a] < 0 ? -1.0 : 0.0);
        hh[a] = std::fabs(om[a]) / h[a];
    }
    double wa[3][2];
    const double* wl[3] = {w_in[0], w_in[1], w_in[2]};
    const double* const* w = w_in;
    if (adapt) {  // adaptive.jl:250-300: AWD0 for orders (1, 1, 1) only
        if (cl.O[0] != 1 || cl.O[1] != 1 || cl.O[2] != 1)
            fail("the adaptive closure in 3D is provided for orders (1, 1, 1) only");
        const double f[3] = {fx[0], fy[0], fz[0]};
        double* wp[3] = {wa[0], wa[1], wa[2]};
        awd0_weights(3, hh, f, Qn[0], St, wp);
        for (int a = 0; a < 3; ++a) wl[a] = wa[a];
        w = wl;
    }
    auto C = [](int i) { return std::sqrt(2.0 * i + 1.0); };  // 0-based
    auto spow = [](double v, int e) { double r = 1.0; for (int k = 0; k < e; ++k) r *= v; return r; };
    const int Nm = cl.Nm;
    vector<double> S((size_t)Nm * Nm, 0.0), Q(Nm, 0.0);
    for (int ax = 0; ax < cl.O[0]; ++ax)
        for (int ay = 0; ay < cl.O[1]; ++ay)
            for (int az = 0; az < cl.O[2]; ++az) {
                if (!cl.valid(ax, ay, az)) continue;
                const int i = cl.idx(ax, ay, az);
                const int ia[3] = {ax, ay, az};
                for (int bx = 0; bx < cl.O[0]; ++bx)
                    for (int by = 0; by < cl.O[1]; ++by)
                        for (int bz = 0; bz < cl.O[2]; ++bz) {
                            if (!cl.valid(bx, by, bz)) continue;
                            const int j = cl.idx(bx, by, bz);
                            const int jb[3] = {bx, by, bz};
                            double& e = S[(size_t)i * Nm + j];
                            if (i == j) e += St;
                            for (int d = 0; d < 3; ++d) {
                                const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
                                if (ia[o1] != jb[o1] || ia[o2] != jb[o2]) continue;
                                const int u = ia[d], v = jb[d];
                                if (u >= v + 1) e -= C(u) * hh[d] * s[d] * C(v) * (1.0 - spow(-1.0, u - v));
                                e += C(u) * hh[d] * spow(s[d], u) * C(v) * spow(s[d], v) * w[d][v + 1];
                            }
                        }
            }
    for (int bx = 0; bx < cl.O[0]; ++bx)
        for (int by = 0; by < cl.O[1]; ++by)
            for (int bz = 0; bz < cl.O[2]; ++bz) {
                if (!cl.valid(bx, by, bz)) continue;
                const int j = cl.idx(bx, by, bz);
                Q[j] = Qn[j];
                Q[j] -= C(bx) * hh[0] * (spow(s[0], bx) * w[0][0] - spow(-s[0], bx)) * fx[cl.fx(by, bz)];
                Q[j] -= C(by) * hh[1] * (spow(s[1], by) * w[1][0] - spow(-s[1], by)) * fy[cl.fy(bx, bz)];
                Q[j] -= C(bz) * hh[2] * (spow(s[2], bz) * w[2][0] - spow(-s[2], bz)) * fz[cl.fz(bx, by)];
            }
    vector<double> x = lu_solve(S, Q, Nm);
    for (int j = 0; j < Nm; ++j) psi[j] = x[j];
    for (int bx = 0; bx < cl.O[0]; ++bx)
        for (int by = 0; by < cl.O[1]; ++by)
            for (int bz = 0; bz < cl.O[2]; ++bz) {
                if (!cl.valid(bx, by, bz)) continue;
                const int j = cl.idx(bx, by, bz);
                double& Fx = fx[cl.fx(by, bz)];
                double& Fy = fy[cl.fy(bx, bz)];
                double& Fz = fz[cl.fz(bx, by)];
                if (bx == 0) Fx = w[0][0] * Fx;
                if (by == 0) Fy = w[1][0] * Fy;
                if (bz == 0) Fz = w[2][0] * Fz;
                Fx += C(bx) * spow(s[0], bx) * w[0][bx + 1] * x[j];
                Fy += C(by) * spow(s[1], by) * w[1][by + 1] * x[j];
                Fz += C(bz) * spow(s[2], bz) * w[2][bz + 1] * x[j];
            }
}

// ½ ∫_{-1}^{1} P_a P_b P_c dx (the 𝒲 constants of scheme_weights.jl:411-412, integrals.jl:254),
// from the closed form of the Wigner 3j symbol with zero projections:
//   (a b c; 0 0 0)² = (J-2a)!(J-2b)!(J-2c)!/(J+1)! · [g!/((g-a)!(g-b)!(g-c)!)]²,  J = a+b+c = 2g,
// zero for odd J or outside the triangle |a-b| <= c <= a+b.
double legendre_triple(int a, int b, int c) {
    const int J = a + b + c;
    if (J % 2 || c > a + b || a > b + c || b > a + c) return 0.0;
    const int g = J / 2;
    auto lf = [](int n) { return std::lgamma(n + 1.0); };
    const double l1 = lf(J - 2 * a) + lf(J - 2 * b) + lf(J - 2 * c) - lf(J + 1);
    const double l2 = lf(g) - lf(g - a) - lf(g - b) - lf(g - c);
    return std::exp(l1 + 2.0 * l2);
}

// The in-group stopping-power moments of sn_flux.jl:92-96 from S⁻/ΔE and S⁺/ΔE (the stopping
// powers at the group's upper and lower energy): (S⁻+S⁺)/2, and (S⁻-S⁺)/(2√3) for Ke > 1.
void stopping_moments(double Sm, double Sp, int Ke, double* Sv) {
    for (int k = 0; k < Ke; ++k) Sv[k] = 0.0;
    Sv[0] = (Sm + Sp) / 2.0;
    if (Ke > 1) Sv[1] = (Sm - Sp) / (2.0 * std::sqrt(3.0));
}

// flux_1D_BFP / flux_2D_BFP / flux_3D_BFP (sn_1D_bfp.jl:44-125, sn_2D_bfp.jl:43-154,
// sn_3D_bfp.jl:51-191; non-adaptive) written once over the axes (E, x, y, z) with orders
// O = (OE, Ox, Oy, Oz); the axes a problem lacks have order 1 and no streaming.  A cell moment is
// a multi-index (e, a, b, c): with full coupling at e + OE*(a + Ox*(b + Oy*c)) (energy fastest),
// without it only one index is non-zero, at that index plus the (O-1) of the axes before it.  The
// face of an axis carries the moments over the other axes by the same rule.
struct Cell4 {
    int O[4];
    bool fc;
    int Nm, F[4];  // cell moments; face moments per axis (F[0]: the energy face)
    Cell4() : O{1, 1, 1, 1}, fc(true), Nm(1), F{1, 1, 1, 1} {}
    Cell4(int oe, int ox, int oy, int oz, bool full) : O{oe, ox, oy, oz}, fc(full) {
        Nm = fc ? oe * ox * oy * oz : oe + ox + oy + oz - 3;
        for (int a = 0; a < 4; ++a) {
            int prod = 1, sum = 0;
            for (int k = 0; k < 4; ++k)
                if (k != a) { prod *= O[k]; sum += O[k] - 1; }
            F[a] = fc ? prod : 1 + sum;
        }
    }
    bool valid(const int* m) const { return fc || (m[0] > 0) + (m[1] > 0) + (m[2] > 0) + (m[3] > 0) <= 1; }
    // index over the axes of `m` except `skip` (-1: all four)
    int index(const int* m, int skip) const {
        int r = 0, stride = 1, off = 0;
        for (int k = 0; k < 4; ++k) {
            if (k == skip) continue;
            if (fc) { r += m[k] * stride; stride *= O[k]; }
            else { if (m[k] > 0) r = m[k] + off; off += O[k] - 1; }
        }
        return r;
    }
};

// One cell of the slowing-down closure.  om: direction cosines (μ, η, ξ), zero on the axes the
// problem lacks; h: cell widths; Sm, Sp: S⁻/ΔE, S⁺/ΔE of the cell's material; Sv[OE]: the
// in-group stopping-power moments / ΔE (sn_flux.jl:92-96); W[OE³]: legendre_triple; f[0..3]: the
// incoming energy, x, y and z face moments, overwritten with the outgoing ones; w[0..3]: the
// closure weights of each axis (scheme_weights_1d).  tang1d: the 1D tangential guard
// (|μ| <= 1e-10: no streaming, the x face passes through, sn_1D_bfp.jl:52-54).
// flux_1D_BFP (sn_1D_bfp.jl:39-125) of orders (2, 2) in (x, E) with full coupling and the coupled
// linear AWD1: adaptive.jl's 2D closure on the axes (x, E) with h = (hx, 1/ΔE), signs (sx, -1), the x
// face and the energy face as inflows and the cell's source moments in their own order (energy
// fastest, as the reference hands them over); then the cell with weight tensors wx[n][jE][iE],
// wE[n][jx][ix] and the in-group stopping-power term.  Moment (ix, iE) at 2 ix + iE; f[1] the x face
// over the energy moments, f[0] the energy face over the x moments.
void flux_bfp1d_awd1(double mu, double dx, double St, double Sm, double Sp, const double* Sv, const double* W,
                     const double* Qn, double* fE, double* fx, double inv_dE, double* psi) {
    const bool tang = std::fabs(mu) <= 1e-10;
    const double sx = tang ? 1.0 : (mu > 0 ? 1.0 : -1.0), hx = tang ? 0.0 : std::fabs(mu) / dx;
    double wx[3][2][2], wE[3][2][2];
    if (tang) {  // no adaptation: DG
        for (int n = 0; n < 3; ++n)
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) wx[n][a][b] = wE[n][a][b] = (a == b && n > 0) ? 1.0 : 0.0;
    } else {
        awd1_weights_2d(hx, inv_dE, sx, -1.0, fx, fE, Qn, wx, wE);
    }
    auto C = [](int i) { return std::sqrt(2.0 * i - 1.0); };  // 1-based
    auto spow = [](double v, int e) { double r = 1.0; for (int k = 0; k < e; ++k) r *= v; return r; };
    const int O = 2, Nm = 4;
    vector<double> S((size_t)Nm * Nm, 0.0), Q(Nm, 0.0);
    for (int ix = 1; ix <= O; ++ix)
        for (int jx = 1; jx <= O; ++jx)
            for (int iE = 1; iE <= O; ++iE)
                for (int jE = 1; jE <= O; ++jE) {
                    const int i = O * (ix - 1) + iE - 1, j = O * (jx - 1) + jE - 1;
                    double& e = S[(size_t)i * Nm + j];
                    if (i == j) e += St;
                    if (iE == jE && ix >= jx + 1) e -= C(ix) * hx * sx * C(jx) * (1.0 - spow(-1.0, ix - jx));
                    e += C(ix) * hx * spow(sx, ix - 1) * C(jx) * spow(sx, jx - 1) * wx[jx][jE - 1][iE - 1];
                    if (ix == jx)
                        for (int kE = 1; kE <= iE - 1; ++kE)
                            for (int wv = 1; wv <= O; ++wv)
                                e += C(iE) * C(jE) * C(kE) * C(wv) * (1.0 - spow(-1.0, iE - kE)) * Sv[wv - 1] *
                                     W[(jE - 1) + O * ((kE - 1) + O * (wv - 1))];
                    e += C(iE) * Sp * spow(-1.0, iE - 1) * C(jE) * spow(-1.0, jE - 1) * wE[jE][jx - 1][ix - 1];
                }
    for (int jx = 1; jx <= O; ++jx)
        for (int jE = 1; jE <= O; ++jE) {
            const int j = O * (jx - 1) + jE - 1;
            Q[j] += Qn[j];
            Q[j] -= C(jx) * hx * (spow(sx, jx - 1) * wx[0][jE - 1][jE - 1] - spow(-sx, jx - 1)) * fx[jE - 1];
            Q[j] -= C(jE) * (spow(-1.0, jE - 1) * Sp * wE[0][jx - 1][jx - 1] - Sm) * fE[jx - 1];
        }
    vector<double> x = lu_solve(S, Q, Nm);
    for (int j = 0; j < Nm; ++j) psi[j] = x[j];
    for (int jx = 1; jx <= O; ++jx)
        for (int jE = 1; jE <= O; ++jE) {
            const int j = O * (jx - 1) + jE - 1;
            if (!tang) {
                if (jx == 1) fx[jE - 1] = wx[0][jE - 1][jE - 1] * fx[jE - 1];
                for (int iE = 1; iE <= O; ++iE) fx[jE - 1] += C(jx) * spow(sx, jx - 1) * wx[jx][jE - 1][iE - 1] * x[j];
            }
            if (jE == 1) fE[jx - 1] = wE[0][jx - 1][jx - 1] * fE[jx - 1];
            for (int ix = 1; ix <= O; ++ix) fE[jx - 1] += C(jE) * spow(-1.0, jE - 1) * wE[jE][jx - 1][ix - 1] * x[j];
        }
}

// adapt: the adaptive closure (AWD0, order 1 on every axis; adaptive.jl on the axes (x[, y[, z]], E)
// as sn_{1,2,3}D_bfp.jl call it, with 1/ΔE as the energy axis' h and the energy face as its inflow),
// not for a tangential 1D direction (sn_1D_bfp.jl:52); inv_dE = 1/ΔE of the group.
void flux_bfp(const double* om, const double* h, double St, double Sm, double Sp, const double* Sv,
              const double* W, const double* Qn, double* const* f, const Cell4& cl, const double* const* w_in,
              bool tang1d, double* psi, bool adapt = false, double inv_dE = 0.0) {
    double s[4] = {-1.0, 0, 0, 0}, hh[4] = {0, 0, 0, 0};
    for (int a = 1; a < 4; ++a) {
        s[a] = om[a - 1] > 0 ? 1.0 : (om[a - 1] < 0 ? -1.0 : 0.0);
        hh[a] = h[a - 1] > 0 ? std::fabs(om[a - 1]) / h[a - 1] : 0.0;
    }
    if (tang1d) { s[1] = 1.0; hh[1] = 0.0; }
    if (adapt && cl.Nm == 4 && cl.fc && cl.O[0] == 2 && cl.O[1] == 2 && cl.O[2] == 1 && cl.O[3] == 1) {
        flux_bfp1d_awd1(om[0], h[0], St, Sm, Sp, Sv, W, Qn, f[0], f[1], inv_dE, psi);
        return;
    }
    double wa[4][2];
    const double* wl[4] = {w_in[0], w_in[1], w_in[2], w_in[3]};
    const double* const* w = w_in;
    if (adapt && !tang1d) {
        if (cl.Nm != 1)
            fail("the adaptive closure with slowing down is provided for order 1 on every axis, or (2, 2) in (x, E) "
                 "with full coupling in 1D");
        // axes x, y, z, then E; an axis a problem lacks has h = 0 and adds nothing to the estimate
        const double h4[4] = {hh[1], hh[2], hh[3], inv_dE}, f4[4] = {f[1][0], f[2][0], f[3][0], f[0][0]};
        double* wp[4] = {wa[1], wa[2], wa[3], wa[0]};
        awd0_weights(4, h4, f4, Qn[0], St, wp);
        for (int a = 0; a < 4; ++a) wl[a] = wa[a];
        w = wl;
    }
    auto C = [](int i) { return std::sqrt(2.0 * i + 1.0); };  // 0-based
    auto spow = [](double v, int e) { double r = 1.0; for (int k = 0; k < e; ++k) r *= v; return r; };
    const int OE = cl.O[0], Nm = cl.Nm;
    vector<int> mi;  // valid multi-indices in the reference's loop order (x outer ... E inner)
    for (int a = 0; a < cl.O[1]; ++a)
        for (int b = 0; b < cl.O[2]; ++b)
            for (int c = 0; c < cl.O[3]; ++c)
                for (int e = 0; e < OE; ++e) {
                    const int m[4] = {e, a, b, c};
                    if (cl.valid(m)) { mi.push_back(e); mi.push_back(a); mi.push_back(b); mi.push_back(c); }
                }
    const int nv = (int)mi.size() / 4;
    vector<double> S((size_t)Nm * Nm, 0.0), Q(Nm, 0.0);
    for (int u = 0; u < nv; ++u) {
        const int* iu = &mi[4 * u];
        const int i = cl.index(iu, -1);
        for (int v = 0; v < nv; ++v) {
            const int* jv = &mi[4 * v];
            const int j = cl.index(jv, -1);
            double& e = S[(size_t)i * Nm + j];
            if (i == j) e += St;
            for (int d = 1; d < 4; ++d) {  // streaming along x, y, z: moments equal on the other axes
                bool same = true;
                for (int k = 0; k < 4; ++k) if (k != d && iu[k] != jv[k]) same = false;
                if (!same) continue;
                const int p = iu[d], q = jv[d];
                if (p >= q + 1) e -= C(p) * hh[d] * s[d] * C(q) * (1.0 - spow(-1.0, p - q));
                e += C(p) * hh[d] * spow(s[d], p) * C(q) * spow(s[d], q) * w[d][q + 1];
            }
            if (iu[1] == jv[1] && iu[2] == jv[2] && iu[3] == jv[3]) {  // slowing down
                const int p = iu[0], q = jv[0];
                for (int k = 0; k < p; ++k)
                    for (int ww = 0; ww < OE; ++ww)
                        e += C(p) * C(q) * C(k) * C(ww) * (1.0 - spow(-1.0, p - k)) * Sv[ww] *
                             W[q + OE * (k + OE * ww)];
                e += C(p) * Sp * spow(-1.0, p) * C(q) * spow(-1.0, q) * w[0][q + 1];
            }
        }
    }
    for (int v = 0; v < nv; ++v) {
        const int* jv = &mi[4 * v];
        const int j = cl.index(jv, -1);
        Q[j] += Qn[j];
        for (int d = 1; d < 4; ++d)
            Q[j] -= C(jv[d]) * hh[d] * (spow(s[d], jv[d]) * w[d][0] - spow(-s[d], jv[d])) * f[d][cl.index(jv, d)];
        Q[j] -= C(jv[0]) * (spow(-1.0, jv[0]) * Sp * w[0][0] - Sm) * f[0][cl.index(jv, 0)];
    }
    vector<double> x = lu_solve(S, Q, Nm);
    for (int j = 0; j < Nm; ++j) psi[j] = x[j];
    for (int v = 0; v < nv; ++v) {
        const int* jv = &mi[4 * v];
        const int j = cl.index(jv, -1);
        for (int d = 0; d < 4; ++d) {
            if (d == 1 && tang1d) continue;
            double& F = f[d][cl.index(jv, d)];
            if (jv[d] == 0) F = w[d][0] * F;
            F += C(jv[d]) * (d == 0 ? spow(-1.0, jv[0]) : spow(s[d], jv[d])) * w[d][jv[d] + 1] * x[j];
        }
    }
}

struct Angular {
    int Nd = 0, Np = 0, Nps = 0;
    vector<double> mu, eta, xi, w;
    Basis B;
    vector<SurfBasis> S;  // one per face
};

}  // namespace

extern "C" {

struct oracle_problem {
    int32_t ndims, Nx, Ny, Nz;
    const double *dx, *dy, *dz;     // cell widths per axis (Nx, Ny, Nz entries)
    const int64_t* mat;             // [Nx*Ny*Nz], 1-based material index, x fastest
    int32_t Nmat, G;
    const double* sigt;             // Radiant Σtot[G, Nmat] column-major
    const double* sigs;             // Radiant Σs[Nmat, G(from), G(to), Lxs+1] column-major
    int32_t Lxs;
    const double* qvol;             // volume source, moment p=1: [G, Nx*Ny*Nz] column-major
    const int32_t* bc;              // [2*ndims] 0 void, 1 reflective, 2 periodic
    const char* quad_type;
    int32_t quad_order, qdims, legendre_order;
    const char* sn_type;
    const char* accel;
    int32_t restart, depth, maxit;
    double tol;
    int32_t nthreads;               // 1 = serial (Radiant's execution model); >1 = OpenMP over directions
    int32_t verbose;
    // surface sources (surface_source.jl): per face, NULL or [G][Np_source][npos_f], position fastest
    const double* ssrc[6];
    int32_t Np_source;
    // continuous slowing down (SN solver type "BCSD", DD of order 1 in energy): 0 = off; Sb the
    // stopping powers at the group boundaries [G+1, Nmat] (column-major), dE the group widths [G]
    int32_t csd;
    const double* Sb;
    const double* dE;
    // SN solver type (SN.jl get_solver_type: 1 BTE, 2 BFP, 3 BCSD, 4 FP, 5 CSD, 6 BFP-EF; 0 = BTE,
    // or BCSD when csd is set); types 2-6 carry the slowing-down term (csd = 1).  T: momentum
    // transfer [G, Nmat] (types 2, 4), siga: absorption [G, Nmat] (types 4, 5), fokker_planck: the
    // angular Fokker-Planck discretisation ("finite-difference" when NULL, "galerkin",
    // "differential-quadrature").
    int32_t solver_type;
    const double* T;
    const double* siga;
    const char* fokker_planck;
    // spatial closure along x (SN.set_scheme("x", scheme, order)): "DD" when NULL, "DG", "DG-",
    // "DG+"; order 0 or 1 = first order.  Orders above 1 and closures other than DD: 1D BTE only.
    const char* scheme;
    int32_t scheme_order;
    // the y axis of 2D problems likewise, and the coupling of the high-order moments
    // (SN.set_is_full_coupling: 0 = full, the default; 1 = Ox + Oy - 1 moments per cell)
    const char* scheme_y;
    int32_t scheme_order_y;
    int32_t partial_coupling;
    const char* scheme_z;       // the z axis of 3D problems
    int32_t scheme_order_z;
    // A rule given point by point (quad_n > 0): the Lebedev rules above order 11, generated from
    // their published orbit parameters by the caller and pinned there by exactness; the
    // upper-hemisphere fold of quadrature.jl:36-54 still applies.
    const double *quad_mu, *quad_eta, *quad_xi, *quad_w;
    int32_t quad_n;
    // the energy closure of the slowing-down types (SN.set_scheme("E", scheme, order)): "DD" when
    // NULL, or "DG"; order 0 or 1 = first order
    const char* scheme_e;
    int32_t scheme_order_e;
    // an external magnetic field (Electromagnetic_Field; sn_flux.jl:116-120): em = 1 adds ℳ_EM Φ to
    // the source (sn_inner_pass.jl:26-32), with the field em_B [3] (x, y, z), the particle charge
    // and the energy boundaries Eb [G+1] (MeV, descending)
    int32_t em;
    double em_B[3];
    double em_charge;
    const double* Eb;
};

struct oracle_stats {
    int64_t niter;
    double resid;
    double rho;
    int32_t converged;
    int32_t passes;
};

}  // extern "C"

namespace {

struct Solver {
    const oracle_problem& P;
    int ndims, Nx, Ny, Nz, Nc, G, Nmat;
    Angular A;
    int Ls;                     // max(pl): Σs sized to Ls+1 (sn_flux.jl:79-83)
    vector<double> sigs;        // [Nmat][G][G][Ls+1] padded copy: sigs[m + Nmat*(gi + G*(gf + G*l))]
    vector<int> matc;           // 0-based material per cell
    int solver = 1;             // SN solver type (1 BTE .. 6 BFP-EF)
    vector<double> sigt_eff;    // Σtot as sn_flux.jl:69-112 prepares it [g + G*m]
    vector<double> Mfp;         // Fokker-Planck ℳ [Np x Np] column-major (types 2, 4)
    double lam0 = 0.0;          // its λ₀
    vector<double> Mem;         // magnetic-field ℳ_EM [G][Np x Np] (empty without a field)
    int K = 1;                  // spatial moments per cell (1D scheme order; 1 elsewhere)
    bool ho = false;            // the general 1D cell solve (order > 1 or a closure other than DD)
    bool adapt = false;         // the adaptive weighted diamond closure (scheme "AWD")
    vector<double> omx, omy;    // its ω per axis
    bool ho2 = false;           // the general 2D cell solve
    Cell2 cl2{1, 1, 1, true};
    bool ho3 = false;           // the general 3D cell solve
    Cell3 cl3{1, 1, 1, true};
    vector<double> omz;
    int Nm1 = 1, Nm2 = 1, Nm3 = 1;  // moments of the x-, y- and z-face fluxes
    bool bfp = false;           // the general slowing-down cell (any order in space or energy)
    Cell4 cl4;
    vector<double> ome, Wt;     // energy closure weights, legendre_triple table [OE³]
    int NmE = 1;                // moments of the energy-face flux (over the spatial axes)
    vector<double> Svec;        // in-group stopping-power moments / ΔE [OE * Nmat]
    int Ncol = 0;               // moment columns: K * Nc (column = spatial moment + K * cell)
    const double* cur_src[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // this pass's surface sources
    // continuous slowing down: incoming / outgoing energy-face angular flux [is + NmE*(c + Nc*n)]
    // (𝚽E12 of sn_flux.jl:166), the current group's S⁻ and S⁺ per material (stopping powers / ΔE)
    vector<double> psiE_in, psiE_out, Sminus, Splus;
    const double* curE = nullptr;  // psiE_in in inhomogeneous passes, null (zero) in homogeneous ones
    // state layout: [Φ (Np*Nc, p fastest) | x-bnd | y-bnd | z-bnd]
    size_t nphi, nbx, nby, nbz, S;
    vector<int> nxm, nxp, nym, nyp, nzm, nzp;  // n_to_n+ per face (1-based or 0)

    explicit Solver(const oracle_problem& p) : P(p) {
        ndims = p.ndims;
        if (ndims < 1 || ndims > 3) fail("Dimension is not 1, 2 or 3.");
        Nx = p.Nx; Ny = ndims >= 2 ? p.Ny : 1; Nz = ndims >= 3 ? p.Nz : 1;
        Nc = Nx * Ny * Nz;
        G = p.G; Nmat = p.Nmat;
        // sn_flux.jl:38-48 -- quadrature and angular operators
        string qt = p.quad_type;
        int Qdims;
        if (qt == "gauss-legendre" || qt == "gauss-lobatto") {  // SN.jl:638-659
            Qdims = 1;
        } else {
            Qdims = (ndims == 3) ? 3 : (p.qdims == 2 ? 2 : 3);
        }
        if (p.quad_n > 0) {
            A.mu.assign(p.quad_mu, p.quad_mu + p.quad_n);
            A.eta.assign(p.quad_eta, p.quad_eta + p.quad_n);
            A.xi.assign(p.quad_xi, p.quad_xi + p.quad_n);
            A.w.assign(p.quad_w, p.quad_w + p.quad_n);
            if (Qdims == 2) fold_upper(A.mu, A.eta, A.xi, A.w);
        } else {
            quadrature(p.quad_order, qt, ndims, Qdims, A.mu, A.eta, A.xi, A.w);
        }
        A.Nd = (int)A.w.size();
        A.B = angular_polynomial_basis(A.mu, A.eta, A.xi, A.w, p.legendre_order, p.sn_type, Qdims);
        A.Np = A.B.Np;
        int nf = 2 * ndims;
        A.S.resize(nf);
        for (int f = 0; f < nf; ++f) {
            A.S[f] = surface_basis(A.mu, A.eta, A.xi, A.w, p.legendre_order, p.sn_type, Qdims, f);
        }
        A.Nps = A.S[nf - 1].Np;  // the reference keeps the last face's count (angular_polynomial_basis.jl:227)
        for (int f = 0; f < nf; ++f)
            if (A.S[f].Np != A.Nps) fail("faces with different half-range basis sizes are not supported");
        auto conv = [&](int f, vector<int>& v) {
            v.assign(A.Nd, 0);
            if (f < nf) for (int n = 0; n < A.Nd; ++n) v[n] = (int)A.S[f].n_to_np[n];
        };
        conv(0, nxm); conv(1, nxp); conv(2, nym); conv(3, nyp); conv(4, nzm); conv(5, nzp);
        // XS (sn_flux.jl:68-83): Σs padded/truncated to Ls = max(pl)
        Ls = 0;
        for (long l : A.B.pl) Ls = std::max(Ls, (int)l);
        sigs.assign((size_t)Nmat * G * G * (Ls + 1), 0.0);
        for (int l = 0; l <= std::min(Ls, p.Lxs); ++l)
            for (int gf = 0; gf < G; ++gf)
                for (int gi = 0; gi < G; ++gi)
                    for (int m = 0; m < Nmat; ++m)
                        sigs[m + (size_t)Nmat * (gi + (size_t)G * (gf + (size_t)G * l))] =
                            p.sigs[m + (size_t)Nmat * (gi + (size_t)G * (gf + (size_t)G * l))];
        matc.resize(Nc);
        for (int c = 0; c < Nc; ++c) {
            long m = p.mat[c];
            if (m < 1 || m > Nmat) fail("material index out of range");
            matc[c] = (int)(m - 1);
        }
        for (int f = 0; f < nf; ++f) {
            if (p.bc[f] < 0 || p.bc[f] > 2) fail("Unknown boundary condition type.");
            if (f == 0 && p.csd) {
                if (!p.Sb || !p.dE) fail("continuous slowing down needs the boundary stopping powers and group widths");
                for (int g = 0; g < p.G; ++g)
                    if (!(p.dE[g] > 0)) fail("energy group widths must be positive");
            }
            if (p.ssrc[f] && (p.Np_source < 0 || p.Np_source > A.Nps))
                fail("surface-source moments exceed the half-range basis of the face");
        }
        // sn_flux.jl:69-112 -- solver type, removal cross section, Fokker-Planck operator
        solver = p.solver_type ? p.solver_type : (p.csd ? 3 : 1);
        if (solver < 1 || solver > 6) fail("Unknown type of solver.");
        if ((solver >= 2) != (p.csd != 0))
            fail("solver types BFP, BCSD, FP, CSD and BFP-EF carry the slowing-down term (csd), BTE does not");
        const double* base = p.sigt;
        if (solver == 4 || solver == 5) {  // absorption removal, no Boltzmann scattering
            if (!p.siga) fail("solver types FP and CSD need the absorption cross sections");
            base = p.siga;
            std::fill(sigs.begin(), sigs.end(), 0.0);
        }
        sigt_eff.assign(base, base + (size_t)G * Nmat);
        if (solver == 2 || solver == 4) {
            if (!p.T) fail("solver types BFP and FP need the momentum transfer");
            const string meth = p.fokker_planck ? p.fokker_planck : "finite-difference";
            // the full-sphere set of the same rule (the Voronoi construction of 2D sets starts there)
            vector<double> S3[4];
            if (p.quad_n > 0) {
                S3[0].assign(p.quad_mu, p.quad_mu + p.quad_n);
                S3[1].assign(p.quad_eta, p.quad_eta + p.quad_n);
                S3[2].assign(p.quad_xi, p.quad_xi + p.quad_n);
            } else if (Qdims >= 2 && meth == "finite-difference" && qt != "gauss-legendre-chebychev") {
                quadrature(p.quad_order, qt, 3, 3, S3[0], S3[1], S3[2], S3[3]);
            }
            Mfp = fokker_planck_matrix(meth, p.quad_order, qt, Qdims, A.mu, A.eta, A.xi, A.w, A.B, S3, lam0);
            for (size_t i = 0; i < sigt_eff.size(); ++i) sigt_eff[i] += p.T[i] * lam0;
        }
        if (p.em) {  // sn_flux.jl:116-120: magnetic fields only, λ₀_EM = 0 leaves Σtot as it is
            if (Qdims != 3) fail("Quadrature on the unit sphere is required for electromagnetic transport.");
            if (!p.Eb) fail("the electromagnetic term needs the energy boundaries");
            Mem = em_matrix(p.em_B, p.em_charge, A.mu, A.eta, A.xi, A.B, G, p.Eb);
        }
        if (solver == 6)  // elastic-free: the self-scattering moment l = 0 leaves the removal
            for (int m = 0; m < Nmat; ++m)
                for (int g = 0; g < G; ++g) sigt_eff[g + (size_t)G * m] -= sig_s(m, g, g, 0);
        // spatial closure (SN.get_schemes, scheme_weights.jl)
        string scheme = p.scheme ? p.scheme : "DD";
        string scheme_y = p.scheme_y ? p.scheme_y : "DD";
        K = p.scheme_order > 1 ? p.scheme_order : 1;
        const int Ky = p.scheme_order_y > 1 ? p.scheme_order_y : 1;
        string scheme_e = p.scheme_e ? p.scheme_e : "DD";
        const int Ke = p.scheme_order_e > 1 ? p.scheme_order_e : 1;
        if (!p.csd && (Ke > 1 || scheme_e != "DD"))
            fail("an energy closure needs the slowing-down term (csd)");
        {  // scheme_weights.jl:32-37: "AWD" on any axis makes every axis adaptive
            const string sz = p.scheme_z ? p.scheme_z : "DD";
            adapt = scheme == "AWD" || (ndims >= 2 && scheme_y == "AWD") || (ndims >= 3 && sz == "AWD") ||
                    (p.csd && scheme_e == "AWD");
            if (adapt) scheme = scheme_y = scheme_e = "AWD";
        }
        if (Ke > 2) fail("energy closures of order above 2 are not supported");
        if (p.csd) {  // scheme_weights.jl with isCSD: axes E, x[, y[, z]], DD or DG each
            const string scheme_z = adapt ? string("AWD") : string(p.scheme_z ? p.scheme_z : "DD");
            const int Kz = p.scheme_order_z > 1 ? p.scheme_order_z : 1;
            const string sy = ndims >= 2 ? scheme_y : (adapt ? "AWD" : "DD"), sz = ndims >= 3 ? scheme_z : (adapt ? "AWD" : "DD");
            const int oy = ndims >= 2 ? Ky : 1, oz = ndims >= 3 ? Kz : 1;
            if (adapt && (Ke > 1 || K > 1 || oy > 1 || oz > 1) &&
                !(ndims == 1 && Ke == 2 && K == 2 && p.partial_coupling == 0))
                fail("the adaptive closure with slowing down is provided for order 1 on every axis, or (2, 2) in "
                     "(x, E) with full coupling in 1D");
            if (Ke > 1 || K > 1 || oy > 1 || oz > 1 || scheme_e != "DD" || scheme != "DD" || sy != "DD" || sz != "DD") {
                for (const string* sc : std::initializer_list<const string*>{&scheme_e, &scheme, &sy, &sz})
                    if (*sc != "DD" && *sc != "DG" && *sc != "AWD") fail("Closure relation not implemented yet.");
                ome = scheme_weights_1d(scheme_e, Ke);
                omx = scheme_weights_1d(scheme, K);
                omy = scheme_weights_1d(sy, oy);
                omz = scheme_weights_1d(sz, oz);
                cl4 = Cell4(Ke, K, oy, oz, p.partial_coupling == 0);
                NmE = cl4.F[0]; Nm1 = cl4.F[1]; Nm2 = cl4.F[2]; Nm3 = cl4.F[3];
                K = cl4.Nm;
                Wt.assign((size_t)Ke * Ke * Ke, 0.0);
                for (int a = 0; a < Ke; ++a)
                    for (int b = 0; b < Ke; ++b)
                        for (int c = 0; c < Ke; ++c) Wt[a + Ke * (b + Ke * c)] = legendre_triple(a, b, c);
                bfp = true;
            }
        } else if (ndims == 1 && (K > 1 || scheme != "DD")) {
            if (adapt && K > 2) fail("the adaptive closure in 1D is provided for orders 1 and 2 only");
            omx = scheme_weights_1d(scheme, K);
            ho = true;
        } else if (ndims == 2 && (K > 1 || Ky > 1 || scheme != "DD" || scheme_y != "DD")) {
            for (const string* sc : {&scheme, &scheme_y})
                if (*sc != "DD" && *sc != "DG" && *sc != "AWD") fail("Closure relation not implemented yet.");
            if (adapt && !((K == 1 && Ky == 1) || (K == 2 && Ky == 2 && p.partial_coupling == 0)))
                fail("the adaptive closure in 2D: orders (1, 1), or (2, 2) with full coupling");
            omx = scheme_weights_1d(scheme, K);
            omy = scheme_weights_1d(scheme_y, Ky);
            const bool fc = p.partial_coupling == 0;
            cl2 = Cell2{K, Ky, fc ? K * Ky : K + Ky - 1, fc};
            Nm1 = Ky;  // x faces carry the y moments (sn_flux.jl get_schemes: Nm[1])
            Nm2 = K;
            K = cl2.Nm;
            ho2 = true;
        } else if (ndims == 3) {
            const string scheme_z = adapt ? string("AWD") : string(p.scheme_z ? p.scheme_z : "DD");
            const int Kz = p.scheme_order_z > 1 ? p.scheme_order_z : 1;
            if (K > 1 || Ky > 1 || Kz > 1 || scheme != "DD" || scheme_y != "DD" || scheme_z != "DD") {
                for (const string* sc : std::initializer_list<const string*>{&scheme, &scheme_y, &scheme_z})
                    if (*sc != "DD" && *sc != "DG" && *sc != "AWD") fail("Closure relation not implemented yet.");
                if (adapt && (K != 1 || Ky != 1 || Kz != 1))
                    fail("the adaptive closure in 3D is provided for orders (1, 1, 1) only");
                omx = scheme_weights_1d(scheme, K);
                omy = scheme_weights_1d(scheme_y, Ky);
                omz = scheme_weights_1d(scheme_z, Kz);
                cl3 = Cell3(K, Ky, Kz, p.partial_coupling == 0);
                Nm1 = cl3.F[0]; Nm2 = cl3.F[1]; Nm3 = cl3.F[2];
                K = cl3.Nm;
                ho3 = true;
            }
        }
        Ncol = K * Nc;
        nphi = (size_t)A.Np * Ncol;
        nbx = (size_t)A.Nps * Nm1 * 2 * Ny * Nz;
        nby = ndims >= 2 ? (size_t)A.Nps * Nm2 * 2 * Nx * Nz : 0;
        nbz = ndims >= 3 ? (size_t)A.Nps * Nm3 * 2 * Nx * Ny : 0;
        S = nphi + nbx + nby + nbz;
    }

    double sig_s(int m, int gi, int gf, int l) const {
        return sigs[m + (size_t)Nmat * (gi + (size_t)G * (gf + (size_t)G * l))];
    }
    double sig_t(int g, int m) const { return sigt_eff[g + (size_t)G * m]; }

    // ----- sweeps --------------------------------------------------------------------
    // DD order-1 cell solve, written once for 1/2/3 axes (sn_1D_bte.jl:150-189,
    // sn_2D_bte.jl:35-117, sn_3D_bte.jl:42-153 with O=1, C=1, w=[-1,2]):
    //   S = ((St + 2hx) + 2hy) + 2hz ;  Q = ((Qn + 2hx*fx) + 2hy*fy) + 2hz*fz ;  psi = Q/S
    //   f_out = -f_in + 2 psi on every axis.

    // sn_sweep_1D.jl:51-124
    void sweep1(int n, int g, double* phi, const double* Ql, const double* bnd_in, double* tmpx) {
        const int Np = A.Np, Nps = A.Nps;
        double mu = A.mu[n];
        const double* Mrow = nullptr;  // row n of M: M[n + Nd*p]
        // incoming face value
        double fx = 0.0;
        int fm = nxm[n], fp = nxp[n];
        const SurfBasis* Sb = nullptr;
        int row = -1;
        if (fm != 0) { Sb = &A.S[0]; row = fm - 1; } else if (fp != 0) { Sb = &A.S[1]; row = fp - 1; }
        bool fwd = mu >= 0;
        int bcin = fwd ? P.bc[0] : P.bc[1];
        // surface source of the entry face (sn_sweep_1D.jl boundary block: the source term first)
        if (Sb && cur_src[fwd ? 0 : 1])
            for (int p = 0; p < P.Np_source; ++p) fx += Sb->M[row + (size_t)Sb->Ndp * p] * cur_src[fwd ? 0 : 1][p];
        if (bcin != 0 && Sb) {
            int src_face = fwd ? (bcin == 1 ? 0 : 1) : (bcin == 1 ? 1 : 0);
            for (int p = 0; p < Nps; ++p)
                fx += Sb->M[row + (size_t)Sb->Ndp * p] * bnd_in[p + (size_t)Nps * src_face];
        }
        (void)Mrow;
        bool tang = std::fabs(mu) <= 1e-10;
        double mue = tang ? 0.0 : mu;
        if (ho) {  // sn_sweep_1D.jl:93-110 with the general cell solve
            vector<double> Qn(K), psi(K);
            for (int k = 0; k < Nx; ++k) {
                int ix = fwd ? k : Nx - 1 - k;
                for (int is = 0; is < K; ++is) {
                    double q = 0.0;
                    for (int p = 0; p < Np; ++p) q += A.B.M[n + (size_t)A.Nd * p] * Ql[p + (size_t)Np * (is + (size_t)K * ix)];
                    Qn[is] = q;
                }
                flux_1d_bte(mu, sig_t(g, matc[ix]), P.dx[ix], Qn.data(), fx, K, omx.data(), psi.data(), adapt);
                for (int is = 0; is < K; ++is)
                    for (int p = 0; p < Np; ++p)
                        phi[p + (size_t)Np * (is + (size_t)K * ix)] += A.B.D[p + (size_t)Np * n] * psi[is];
            }
        } else
        for (int k = 0; k < Nx; ++k) {
            int ix = fwd ? k : Nx - 1 - k;
            int c = ix;
            double Qn = 0.0;
            for (int p = 0; p < Np; ++p) Qn += A.B.M[n + (size_t)A.Nd * p] * Ql[p + (size_t)Np * c];
            double St = sig_t(g, matc[c]);
            double hx = std::fabs(mue) / P.dx[ix];
            double Sm = St + 2.0 * hx;
            double Q = Qn + 2.0 * hx * fx;
            double Ein = 0.0;
            if (P.csd) {  // flux_1D_BFP with O = 1 in space and energy (sn_1D_bfp.jl)
                const int m = matc[c];
                Ein = curE ? curE[(size_t)n * Nc + c] : 0.0;
                Sm = Sm + 2.0 * Splus[m];
                Q = Q + (Splus[m] + Sminus[m]) * Ein;
            }
            double psi = Q / Sm;
            if (P.csd) psiE_out[(size_t)n * Nc + c] = 2.0 * psi - Ein;
            if (!tang) fx = -fx + 2.0 * psi;
            for (int p = 0; p < Np; ++p) phi[p + (size_t)Np * c] += A.B.D[p + (size_t)Np * n] * psi;
        }
        if (Sb) {
            int of = fwd ? 1 : 0;
            for (int p = 0; p < Nps; ++p) tmpx[p + (size_t)Nps * of] += Sb->D[p + (size_t)Nps * row] * fx;
        }
    }

    // sn_sweep_2D.jl:54-200 with the general cell: columns (x) outer, rows (y) inner.  The x-face
    // moments of every row are carried across the columns, the y-face moments along a column;
    // face arrays are b[p + Nps*(is + Nm*(f + 2*pos))], surface sources enter moment 1 only.
    void sweep2_ho(int n, int g, double* phi, const double* Ql, const double* bx_in, const double* by_in, double* tx,
                   double* ty) {
        const int Np = A.Np, Nps = A.Nps, Nd = A.Nd;
        const double mu = A.mu[n], eta = A.eta[n];
        const bool fx_ = mu >= 0, fy_ = eta >= 0;
        const SurfBasis *Sx = nullptr, *Sy = nullptr;
        int rx = -1, ry = -1;
        if (nxm[n]) { Sx = &A.S[0]; rx = nxm[n] - 1; } else if (nxp[n]) { Sx = &A.S[1]; rx = nxp[n] - 1; }
        if (nym[n]) { Sy = &A.S[2]; ry = nym[n] - 1; } else if (nyp[n]) { Sy = &A.S[3]; ry = nyp[n] - 1; }
        auto inflow = [&](int face_bc, bool fwd, const SurfBasis* Sb, int row, const double* b, int Nm, size_t pos,
                          const double* src, size_t npos, double* f) {
            if (src && Sb)
                for (int p = 0; p < P.Np_source; ++p) f[0] += Sb->M[row + (size_t)Sb->Ndp * p] * src[(size_t)p * npos + pos];
            if (face_bc == 0 || !Sb) return;
            const int sf = fwd ? (face_bc == 1 ? 0 : 1) : (face_bc == 1 ? 1 : 0);
            for (int p = 0; p < Nps; ++p)
                for (int is = 0; is < Nm; ++is)
                    f[is] += Sb->M[row + (size_t)Sb->Ndp * p] * b[p + (size_t)Nps * (is + (size_t)Nm * (sf + 2 * pos))];
        };
        vector<double> fxm((size_t)Nm1 * Ny, 0.0), fym(Nm2), Qn(K), psi(K);
        for (int kx = 0; kx < Nx; ++kx) {
            const int ix = fx_ ? kx : Nx - 1 - kx;
            std::fill(fym.begin(), fym.end(), 0.0);
            inflow(fy_ ? P.bc[2] : P.bc[3], fy_, Sy, ry, by_in, Nm2, ix, cur_src[fy_ ? 2 : 3], (size_t)Nx, fym.data());
            for (int ky = 0; ky < Ny; ++ky) {
                const int iy = fy_ ? ky : Ny - 1 - ky;
                double* fx = &fxm[(size_t)Nm1 * iy];
                if (kx == 0)
                    inflow(fx_ ? P.bc[0] : P.bc[1], fx_, Sx, rx, bx_in, Nm1, iy, cur_src[fx_ ? 0 : 1], (size_t)Ny, fx);
                const int c = ix + Nx * iy;
                for (int is = 0; is < K; ++is) {
                    double q = 0.0;
                    for (int p = 0; p < Np; ++p) q += A.B.M[n + (size_t)Nd * p] * Ql[p + (size_t)Np * (is + (size_t)K * c)];
                    Qn[is] = q;
                }
                flux_2d_bte(mu, eta, sig_t(g, matc[c]), P.dx[ix], P.dy[iy], Qn.data(), fx, fym.data(), cl2, omx.data(),
                            omy.data(), psi.data(), adapt);
                for (int is = 0; is < K; ++is)
                    for (int p = 0; p < Np; ++p)
                        phi[p + (size_t)Np * (is + (size_t)K * c)] += A.B.D[p + (size_t)Np * n] * psi[is];
                if (kx == Nx - 1 && Sx)
                    for (int p = 0; p < Nps; ++p)
                        for (int is = 0; is < Nm1; ++is)
                            tx[p + (size_t)Nps * (is + (size_t)Nm1 * ((fx_ ? 1 : 0) + 2 * iy))] +=
                                Sx->D[p + (size_t)Nps * rx] * fx[is];
                if (ky == Ny - 1 && Sy)
                    for (int p = 0; p < Nps; ++p)
                        for (int is = 0; is < Nm2; ++is)
                            ty[p + (size_t)Nps * (is + (size_t)Nm2 * ((fy_ ? 1 : 0) + 2 * ix))] +=
                                Sy->D[p + (size_t)Nps * ry] * fym[is];
            }
        }
    }

    // sn_sweep_3D.jl:57-240 with the general cell: x outer, y, z inner.  x-face moments per (y, z)
    // are carried across the columns, y-face moments per z along a column, z-face moments along a
    // pillar; face arrays b[p + Nps*(m + Nm*(f + 2*pos))], surface sources enter moment 1 only.
    void sweep3_ho(int n, int g, double* phi, const double* Ql, const double* bx_in, const double* by_in,
                   const double* bz_in, double* tx, double* ty, double* tz) {
        const int Np = A.Np, Nps = A.Nps, Nd = A.Nd;
        const double om[3] = {A.mu[n], A.eta[n], A.xi[n]};
        const bool fx_ = om[0] >= 0, fy_ = om[1] >= 0, fz_ = om[2] >= 0;
        const SurfBasis *Sx = nullptr, *Sy = nullptr, *Sz = nullptr;
        int rx = -1, ry = -1, rz = -1;
        if (nxm[n]) { Sx = &A.S[0]; rx = nxm[n] - 1; } else if (nxp[n]) { Sx = &A.S[1]; rx = nxp[n] - 1; }
        if (nym[n]) { Sy = &A.S[2]; ry = nym[n] - 1; } else if (nyp[n]) { Sy = &A.S[3]; ry = nyp[n] - 1; }
        if (nzm[n]) { Sz = &A.S[4]; rz = nzm[n] - 1; } else if (nzp[n]) { Sz = &A.S[5]; rz = nzp[n] - 1; }
        auto add_src = [&](const SurfBasis* Sb, int row, const double* src, size_t npos, size_t pos, double* f) {
            if (src && Sb)
                for (int p = 0; p < P.Np_source; ++p) f[0] += Sb->M[row + (size_t)Sb->Ndp * p] * src[(size_t)p * npos + pos];
        };
        auto add_bc = [&](int face_bc, bool fwd, const SurfBasis* Sb, int row, const double* b, int Nm, size_t pos,
                          double* f) {
            if (face_bc == 0 || !Sb) return;
            const int sf = fwd ? (face_bc == 1 ? 0 : 1) : (face_bc == 1 ? 1 : 0);
            for (int p = 0; p < Nps; ++p)
                for (int is = 0; is < Nm; ++is)
                    f[is] += Sb->M[row + (size_t)Sb->Ndp * p] * b[p + (size_t)Nps * (is + (size_t)Nm * (sf + 2 * pos))];
        };
        const double* w[3] = {omx.data(), omy.data(), omz.data()};
        vector<double> FX((size_t)Nm1 * Ny * Nz, 0.0), FY((size_t)Nm2 * Nz), FZ(Nm3), Qn(K), psi(K);
        for (int kx = 0; kx < Nx; ++kx) {
            const int ix = fx_ ? kx : Nx - 1 - kx;
            std::fill(FY.begin(), FY.end(), 0.0);
            for (int ky = 0; ky < Ny; ++ky) {
                const int iy = fy_ ? ky : Ny - 1 - ky;
                std::fill(FZ.begin(), FZ.end(), 0.0);
                add_src(Sz, rz, cur_src[fz_ ? 4 : 5], (size_t)Nx * Ny, (size_t)ix + (size_t)Nx * iy, FZ.data());
                add_bc(fz_ ? P.bc[4] : P.bc[5], fz_, Sz, rz, bz_in, Nm3, (size_t)ix + (size_t)Nx * iy, FZ.data());
                for (int kz = 0; kz < Nz; ++kz) {
                    const int iz = fz_ ? kz : Nz - 1 - kz;
                    double* fy = &FY[(size_t)Nm2 * iz];
                    double* fx = &FX[(size_t)Nm1 * (iy + (size_t)Ny * iz)];
                    if (ky == 0) {
                        add_src(Sy, ry, cur_src[fy_ ? 2 : 3], (size_t)Nx * Nz, (size_t)ix + (size_t)Nx * iz, fy);
                        add_bc(fy_ ? P.bc[2] : P.bc[3], fy_, Sy, ry, by_in, Nm2, (size_t)ix + (size_t)Nx * iz, fy);
                    }
                    if (kx == 0) {
                        add_src(Sx, rx, cur_src[fx_ ? 0 : 1], (size_t)Ny * Nz, (size_t)iy + (size_t)Ny * iz, fx);
                        add_bc(fx_ ? P.bc[0] : P.bc[1], fx_, Sx, rx, bx_in, Nm1, (size_t)iy + (size_t)Ny * iz, fx);
                    }
                    const int c = ix + Nx * (iy + Ny * iz);
                    for (int is = 0; is < K; ++is) {
                        double q = 0.0;
                        for (int p = 0; p < Np; ++p) q += A.B.M[n + (size_t)Nd * p] * Ql[p + (size_t)Np * (is + (size_t)K * c)];
                        Qn[is] = q;
                    }
                    const double h[3] = {P.dx[ix], P.dy[iy], P.dz[iz]};
                    flux_3d_bte(om, h, sig_t(g, matc[c]), Qn.data(), fx, fy, FZ.data(), cl3, w, psi.data(), adapt);
                    for (int is = 0; is < K; ++is)
                        for (int p = 0; p < Np; ++p)
                            phi[p + (size_t)Np * (is + (size_t)K * c)] += A.B.D[p + (size_t)Np * n] * psi[is];
                    if (kx == Nx - 1 && Sx)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm1; ++is)
                                tx[p + (size_t)Nps * (is + (size_t)Nm1 * ((fx_ ? 1 : 0) + 2 * (iy + (size_t)Ny * iz)))] +=
                                    Sx->D[p + (size_t)Nps * rx] * fx[is];
                    if (ky == Ny - 1 && Sy)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm2; ++is)
                                ty[p + (size_t)Nps * (is + (size_t)Nm2 * ((fy_ ? 1 : 0) + 2 * (ix + (size_t)Nx * iz)))] +=
                                    Sy->D[p + (size_t)Nps * ry] * fy[is];
                    if (kz == Nz - 1 && Sz)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm3; ++is)
                                tz[p + (size_t)Nps * (is + (size_t)Nm3 * ((fz_ ? 1 : 0) + 2 * (ix + (size_t)Nx * iy)))] +=
                                    Sz->D[p + (size_t)Nps * rz] * FZ[is];
                }
            }
        }
    }

    // sn_sweep_1D.jl:51-124, sn_sweep_2D.jl:54-200 and sn_sweep_3D.jl:57-240 with the
    // slowing-down cell of any order (flux_bfp): x outer, y, z inner, as sweep3_ho; the axes a
    // problem lacks have one position.  Face arrays b[p + Nps*(m + Nm*(f + 2*pos))]; surface
    // sources enter face moment 1 only.  The energy face of each cell comes in from psiE (the
    // previous group's outflow, zero in homogeneous passes) and leaves into psiE_out.
    void sweep_bfp(int n, int g, double* phi, const double* Ql, const double* bx_in, const double* by_in,
                   const double* bz_in, double* tx, double* ty, double* tz) {
        const int Np = A.Np, Nps = A.Nps, Nd = A.Nd;
        const double om[3] = {A.mu[n], ndims >= 2 ? A.eta[n] : 0.0, ndims >= 3 ? A.xi[n] : 0.0};
        const bool fx_ = om[0] >= 0, fy_ = om[1] >= 0, fz_ = om[2] >= 0;
        const bool tang = ndims == 1 && std::fabs(om[0]) <= 1e-10;
        const SurfBasis *Sx = nullptr, *Sy = nullptr, *Sz = nullptr;
        int rx = -1, ry = -1, rz = -1;
        if (nxm[n]) { Sx = &A.S[0]; rx = nxm[n] - 1; } else if (nxp[n]) { Sx = &A.S[1]; rx = nxp[n] - 1; }
        if (ndims >= 2) {
            if (nym[n]) { Sy = &A.S[2]; ry = nym[n] - 1; } else if (nyp[n]) { Sy = &A.S[3]; ry = nyp[n] - 1; }
        }
        if (ndims >= 3) {
            if (nzm[n]) { Sz = &A.S[4]; rz = nzm[n] - 1; } else if (nzp[n]) { Sz = &A.S[5]; rz = nzp[n] - 1; }
        }
        auto add_src = [&](const SurfBasis* Sb, int row, const double* src, size_t npos, size_t pos, double* f) {
            if (src && Sb)
                for (int p = 0; p < P.Np_source; ++p) f[0] += Sb->M[row + (size_t)Sb->Ndp * p] * src[(size_t)p * npos + pos];
        };
        auto add_bc = [&](int face_bc, bool fwd, const SurfBasis* Sb, int row, const double* b, int Nm, size_t pos,
                          double* f) {
            if (face_bc == 0 || !Sb) return;
            const int sf = fwd ? (face_bc == 1 ? 0 : 1) : (face_bc == 1 ? 1 : 0);
            for (int p = 0; p < Nps; ++p)
                for (int is = 0; is < Nm; ++is)
                    f[is] += Sb->M[row + (size_t)Sb->Ndp * p] * b[p + (size_t)Nps * (is + (size_t)Nm * (sf + 2 * pos))];
        };
        const double* w[4] = {ome.data(), omx.data(), omy.data(), omz.data()};
        vector<double> FX((size_t)Nm1 * Ny * Nz, 0.0), FY((size_t)Nm2 * Nz), FZ(Nm3), FE(NmE), Qn(K), psi(K);
        for (int kx = 0; kx < Nx; ++kx) {
            const int ix = fx_ ? kx : Nx - 1 - kx;
            std::fill(FY.begin(), FY.end(), 0.0);
            for (int ky = 0; ky < Ny; ++ky) {
                const int iy = fy_ ? ky : Ny - 1 - ky;
                std::fill(FZ.begin(), FZ.end(), 0.0);
                if (ndims >= 3) {
                    add_src(Sz, rz, cur_src[fz_ ? 4 : 5], (size_t)Nx * Ny, (size_t)ix + (size_t)Nx * iy, FZ.data());
                    add_bc(fz_ ? P.bc[4] : P.bc[5], fz_, Sz, rz, bz_in, Nm3, (size_t)ix + (size_t)Nx * iy, FZ.data());
                }
                for (int kz = 0; kz < Nz; ++kz) {
                    const int iz = fz_ ? kz : Nz - 1 - kz;
                    double* fy = &FY[(size_t)Nm2 * iz];
                    double* fx = &FX[(size_t)Nm1 * (iy + (size_t)Ny * iz)];
                    if (ky == 0 && ndims >= 2) {
                        add_src(Sy, ry, cur_src[fy_ ? 2 : 3], (size_t)Nx * Nz, (size_t)ix + (size_t)Nx * iz, fy);
                        add_bc(fy_ ? P.bc[2] : P.bc[3], fy_, Sy, ry, by_in, Nm2, (size_t)ix + (size_t)Nx * iz, fy);
                    }
                    if (kx == 0) {
                        add_src(Sx, rx, cur_src[fx_ ? 0 : 1], (size_t)Ny * Nz, (size_t)iy + (size_t)Ny * iz, fx);
                        add_bc(fx_ ? P.bc[0] : P.bc[1], fx_, Sx, rx, bx_in, Nm1, (size_t)iy + (size_t)Ny * iz, fx);
                    }
                    const int c = ix + Nx * (iy + Ny * iz);
                    for (int is = 0; is < K; ++is) {
                        double q = 0.0;
                        for (int p = 0; p < Np; ++p) q += A.B.M[n + (size_t)Nd * p] * Ql[p + (size_t)Np * (is + (size_t)K * c)];
                        Qn[is] = q;
                    }
                    double* fe_out = &psiE_out[(size_t)NmE * (c + (size_t)Nc * n)];
                    for (int is = 0; is < NmE; ++is) FE[is] = curE ? curE[is + (size_t)NmE * (c + (size_t)Nc * n)] : 0.0;
                    const double h[3] = {P.dx[ix], ndims >= 2 ? P.dy[iy] : 0.0, ndims >= 3 ? P.dz[iz] : 0.0};
                    const int m = matc[c];
                    double* f[4] = {FE.data(), fx, fy, FZ.data()};
                    flux_bfp(om, h, sig_t(g, m), Sminus[m], Splus[m], &Svec[(size_t)cl4.O[0] * m], Wt.data(), Qn.data(),
                             f, cl4, w, tang, psi.data(), adapt, 1.0 / P.dE[g]);
                    for (int is = 0; is < NmE; ++is) fe_out[is] = FE[is];
                    for (int is = 0; is < K; ++is)
                        for (int p = 0; p < Np; ++p)
                            phi[p + (size_t)Np * (is + (size_t)K * c)] += A.B.D[p + (size_t)Np * n] * psi[is];
                    if (kx == Nx - 1 && Sx)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm1; ++is)
                                tx[p + (size_t)Nps * (is + (size_t)Nm1 * ((fx_ ? 1 : 0) + 2 * (iy + (size_t)Ny * iz)))] +=
                                    Sx->D[p + (size_t)Nps * rx] * fx[is];
                    if (ky == Ny - 1 && Sy)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm2; ++is)
                                ty[p + (size_t)Nps * (is + (size_t)Nm2 * ((fy_ ? 1 : 0) + 2 * (ix + (size_t)Nx * iz)))] +=
                                    Sy->D[p + (size_t)Nps * ry] * fy[is];
                    if (kz == Nz - 1 && Sz)
                        for (int p = 0; p < Nps; ++p)
                            for (int is = 0; is < Nm3; ++is)
                                tz[p + (size_t)Nps * (is + (size_t)Nm3 * ((fz_ ? 1 : 0) + 2 * (ix + (size_t)Nx * iy)))] +=
                                    Sz->D[p + (size_t)Nps * rz] * FZ[is];
                }
            }
        }
    }

    // incoming boundary value for a face array b[p + Nps*(f + 2*pos)] (sweep BC branches).
    // src: the entry face's surface source of the current group ([Np_source][npos]) or null; its
    // term comes first, as in sn_sweep_2D.jl:70-99.
    double bc_in(int face_bc, bool fwd, const SurfBasis* Sb, int row, const double* b, size_t pos,
                 const double* src = nullptr, size_t npos = 0) {
        double v = 0.0;
        if (src && Sb)
            for (int p = 0; p < P.Np_source; ++p) v += Sb->M[row + (size_t)Sb->Ndp * p] * src[(size_t)p * npos + pos];
        if (face_bc == 0 || !Sb) return v;
        int src_face = fwd ? (face_bc == 1 ? 0 : 1) : (face_bc == 1 ? 1 : 0);
        for (int p = 0; p < A.Nps; ++p) v += Sb->M[row + (size_t)Sb->Ndp * p] * b[p + (size_t)A.Nps * (src_face + 2 * pos)];
        return v;
    }

    // sn_sweep_2D.jl:54-184 and sn_sweep_3D.jl:57-240 (z collapses to one layer in 2D).
    void sweep23(int n, int g, double* phi, const double* Ql, const double* bx_in, const double* by_in,
                 const double* bz_in, double* tx, double* ty, double* tz) {
        const int Np = A.Np, Nps = A.Nps, Nd = A.Nd;
        double mu = A.mu[n], eta = A.eta[n], xi = A.xi[n];
        bool fx_ = mu >= 0, fy_ = eta >= 0, fz_ = xi >= 0;
        const SurfBasis *Sx = nullptr, *Sy = nullptr, *Sz = nullptr;
        int rx = -1, ry = -1, rz = -1;
        if (nxm[n]) { Sx = &A.S[0]; rx = nxm[n] - 1; } else if (nxp[n]) { Sx = &A.S[1]; rx = nxp[n] - 1; }
        if (nym[n]) { Sy = &A.S[2]; ry = nym[n] - 1; } else if (nyp[n]) { Sy = &A.S[3]; ry = nyp[n] - 1; }
        if (ndims == 3) {
            if (nzm[n]) { Sz = &A.S[4]; rz = nzm[n] - 1; } else if (nzp[n]) { Sz = &A.S[5]; rz = nzp[n] - 1; }
        }
        int bcx = fx_ ? P.bc[0] : P.bc[1];
        int bcy = fy_ ? P.bc[2] : P.bc[3];
        int bcz = ndims == 3 ? (fz_ ? P.bc[4] : P.bc[5]) : 0;
        vector<double> Fx((size_t)Ny * Nz, 0.0);  // x-face flux per (iy,iz)
        vector<double> Fy(Nz, 0.0);                // y-face flux per iz (reset per ix)
        for (int kx = 0; kx < Nx; ++kx) {
            int ix = fx_ ? kx : Nx - 1 - kx;
            std::fill(Fy.begin(), Fy.end(), 0.0);
            for (int ky = 0; ky < Ny; ++ky) {
                int iy = fy_ ? ky : Ny - 1 - ky;
                double Fz = 0.0;
                if (ndims == 3)
                    Fz = bc_in(bcz, fz_, Sz, rz, bz_in, (size_t)ix + (size_t)Nx * iy, cur_src[fz_ ? 4 : 5],
                               (size_t)Nx * Ny);
                for (int kz = 0; kz < Nz; ++kz) {
                    int iz = fz_ ? kz : Nz - 1 - kz;
                    if (ky == 0)
                        Fy[iz] = bc_in(bcy, fy_, Sy, ry, by_in, (size_t)ix + (size_t)Nx * iz, cur_src[fy_ ? 2 : 3],
                                       (size_t)Nx * Nz);
                    if (kx == 0)
                        Fx[iy + (size_t)Ny * iz] = bc_in(bcx, fx_, Sx, rx, bx_in, (size_t)iy + (size_t)Ny * iz,
                                                         cur_src[fx_ ? 0 : 1], (size_t)Ny * Nz);
                    size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
                    double Qn = 0.0;
                    for (int p = 0; p < Np; ++p) Qn += A.B.M[n + (size_t)Nd * p] * Ql[p + (size_t)Np * c];
                    double St = sig_t(g, matc[c]);
                    double hx = std::fabs(mu) / P.dx[ix];
                    double hy = std::fabs(eta) / P.dy[iy];
                    double& fxr = Fx[iy + (size_t)Ny * iz];
                    double& fyr = Fy[iz];
                    double psi;
                    // flux_2D_BFP / flux_3D_BFP with O = 1: the energy terms come after the spatial ones
                    const double Ein = (P.csd && curE) ? curE[(size_t)n * Nc + c] : 0.0;
                    if (ndims == 2) {
                        double Sm = St + 2.0 * hx;
                        Sm = Sm + 2.0 * hy;
                        double Q = Qn + 2.0 * hx * fxr;
                        Q = Q + 2.0 * hy * fyr;
                        if (P.csd) {
                            Sm = Sm + 2.0 * Splus[matc[c]];
                            Q = Q + (Splus[matc[c]] + Sminus[matc[c]]) * Ein;
                        }
                        psi = Q / Sm;
                    } else {
                        double hz = std::fabs(xi) / P.dz[iz];
                        double Sm = ((St + 2.0 * hx) + 2.0 * hy) + 2.0 * hz;
                        double Q = ((Qn + 2.0 * hx * fxr) + 2.0 * hy * fyr) + 2.0 * hz * Fz;
                        if (P.csd) {
                            Sm = Sm + 2.0 * Splus[matc[c]];
                            Q = Q + (Splus[matc[c]] + Sminus[matc[c]]) * Ein;
                        }
                        psi = Q / Sm;
                        Fz = -Fz + 2.0 * psi;
                    }
                    if (P.csd) psiE_out[(size_t)n * Nc + c] = 2.0 * psi - Ein;
                    fxr = -fxr + 2.0 * psi;
                    fyr = -fyr + 2.0 * psi;
                    for (int p = 0; p < Np; ++p) phi[p + (size_t)Np * c] += A.B.D[p + (size_t)Np * n] * psi;
                    if (kx == Nx - 1 && Sx) {
                        int of = fx_ ? 1 : 0;
                        for (int p = 0; p < Nps; ++p)
                            tx[p + (size_t)Nps * (of + 2 * ((size_t)iy + (size_t)Ny * iz))] += Sx->D[p + (size_t)Nps * rx] * fxr;
                    }
                    if (ky == Ny - 1 && Sy) {
                        int of = fy_ ? 1 : 0;
                        for (int p = 0; p < Nps; ++p)
                            ty[p + (size_t)Nps * (of + 2 * ((size_t)ix + (size_t)Nx * iz))] += Sy->D[p + (size_t)Nps * ry] * fyr;
                    }
                    if (ndims == 3 && kz == Nz - 1 && Sz) {
                        int of = fz_ ? 1 : 0;
                        for (int p = 0; p < Nps; ++p)
                            tz[p + (size_t)Nps * (of + 2 * ((size_t)ix + (size_t)Nx * iy))] += Sz->D[p + (size_t)Nps * rz] * Fz;
                    }
                }
            }
        }
    }

    // sn_inner_pass.jl:212-332 -- one pass T(z) (homogeneous=false) or A z (true), in place on
    // the state z = [Φ, x-bnd, y-bnd, z-bnd].  With nthreads>1 the direction loop is split over
    // OpenMP threads with private accumulators, reduced in thread order.
    void pass(int g, bool homog, vector<double>& z, const vector<double>& Qlout, vector<double>& Ql) {
        const int Np = A.Np;
        double* phi = z.data();
        double* bx = phi + nphi;
        // surface sources enter the inhomogeneous passes only (sn_inner_pass.jl: sources with
        // homogeneous = false)
        const size_t npf[6] = {(size_t)Ny * Nz, (size_t)Ny * Nz, (size_t)Nx * Nz, (size_t)Nx * Nz,
                               (size_t)Nx * Ny, (size_t)Nx * Ny};
        for (int f = 0; f < 6; ++f)
            cur_src[f] = (!homog && f < 2 * ndims && P.ssrc[f]) ? P.ssrc[f] + (size_t)g * P.Np_source * npf[f] : nullptr;
        if (P.csd) {  // sn_flux.jl:183-188 (S / ΔE); the energy inflow is a fixed source (sn_inner_pass.jl:38)
            Sminus.assign(Nmat, 0.0);
            Splus.assign(Nmat, 0.0);
            for (int m = 0; m < Nmat; ++m) {
                Sminus[m] = P.Sb[g + (size_t)(G + 1) * m] / P.dE[g];
                Splus[m] = P.Sb[(g + 1) + (size_t)(G + 1) * m] / P.dE[g];
            }
            if (bfp) {  // sn_flux.jl:92-96: S = [(S⁻+S⁺)/2, (S⁻-S⁺)/(2√3), 0...] / ΔE
                const int Ke = cl4.O[0];
                Svec.assign((size_t)Ke * Nmat, 0.0);
                for (int m = 0; m < Nmat; ++m) stopping_moments(Sminus[m], Splus[m], Ke, &Svec[(size_t)Ke * m]);
            }
            const size_t ne = (size_t)A.Nd * Nc * NmE;
            if (psiE_in.size() != ne) psiE_in.assign(ne, 0.0);
            psiE_out.assign(ne, 0.0);
            curE = homog ? nullptr : psiE_in.data();
        }
        // Ql = (homog ? 0 : Qlout) + Σs_l Φl   (scattering_source.jl:25-30; sn_inner_pass.jl:22
        // leaves the in-group scattering out for solver types 4, 5 and 6)
        const bool within = solver != 4 && solver != 5 && solver != 6;
        for (int c = 0; c < Ncol; ++c) {  // moment columns (spatial moment + K * cell)
            int m = matc[c / K];
            for (int p = 0; p < Np; ++p) {
                double q = homog ? 0.0 : Qlout[p + (size_t)Np * c];
                if (within) q += sig_s(m, g, g, (int)A.B.pl[p]) * phi[p + (size_t)Np * c];
                Ql[p + (size_t)Np * c] = q;
            }
        }
        // Ql += T ℳ Φl   (fokker_planck_source.jl:26-36, solver types 2 and 4)
        if (solver == 2 || solver == 4) {
            for (int c = 0; c < Ncol; ++c) {
                const double Tm = P.T[g + (size_t)G * matc[c / K]];
                for (int n = 0; n < Np; ++n)
                    for (int k = 0; k < Np; ++k)
                        Ql[n + (size_t)Np * c] += Tm * Mfp[n + (size_t)Np * k] * phi[k + (size_t)Np * c];
            }
        }
        // Ql += ℳ_EM Φl   (sn_inner_pass.jl:26-32)
        if (!Mem.empty()) {
            const double* Me = &Mem[(size_t)g * Np * Np];
            for (int c = 0; c < Ncol; ++c)
                for (int n = 0; n < Np; ++n)
                    for (int k = 0; k < Np; ++k)
                        Ql[n + (size_t)Np * c] += Me[n + (size_t)Np * k] * phi[k + (size_t)Np * c];
        }
        vector<double> bin(bx, bx + nbx + nby + nbz);
        int nt = std::max(1, P.nthreads);
        if (nt == 1) {
            std::fill(phi, phi + nphi, 0.0);
            vector<double> tmp(nbx + nby + nbz, 0.0);
            for (int n = 0; n < A.Nd; ++n) direction(n, g, phi, Ql.data(), bin.data(), tmp.data());
            std::copy(tmp.begin(), tmp.end(), bx);
        } else {
#ifdef _OPENMP
            vector<vector<double>> lphi(nt), ltmp(nt);
#pragma omp parallel num_threads(nt)
            {
                int t = omp_get_thread_num();
                lphi[t].assign(nphi, 0.0);
                ltmp[t].assign(nbx + nby + nbz, 0.0);
#pragma omp for schedule(static)
                for (int n = 0; n < A.Nd; ++n) direction(n, g, lphi[t].data(), Ql.data(), bin.data(), ltmp[t].data());
            }
            std::fill(phi, phi + nphi, 0.0);
            std::fill(bx, bx + nbx + nby + nbz, 0.0);
            for (int t = 0; t < nt; ++t) {
                for (size_t i = 0; i < nphi; ++i) phi[i] += lphi[t][i];
                for (size_t i = 0; i < nbx + nby + nbz; ++i) bx[i] += ltmp[t][i];
            }
#else
            fail("oracle built without OpenMP");
#endif
        }
    }

    void direction(int n, int g, double* phi, const double* Ql, const double* bin, double* tmp) {
        if (bfp) sweep_bfp(n, g, phi, Ql, bin, bin + nbx, bin + nbx + nby, tmp, tmp + nbx, tmp + nbx + nby);
        else if (ndims == 1) sweep1(n, g, phi, Ql, bin, tmp);
        else if (ho2) sweep2_ho(n, g, phi, Ql, bin, bin + nbx, tmp, tmp + nbx);
        else if (ho3) sweep3_ho(n, g, phi, Ql, bin, bin + nbx, bin + nbx + nby, tmp, tmp + nbx, tmp + nbx + nby);
        else sweep23(n, g, phi, Ql, bin, bin + nbx, bin + nbx + nby, tmp, tmp + nbx, tmp + nbx + nby);
    }

    // ----- Krylov state algebra (tools/krylov_state.jl:28-144) ------------------------
    double dot(const vector<double>& a, const vector<double>& b) const {
        size_t off[5] = {0, nphi, nphi + nbx, nphi + nbx + nby, S};
        double s = 0.0;
        for (int k = 0; k < 4; ++k) {
            double sk = 0.0;
            for (size_t i = off[k]; i < off[k + 1]; ++i) sk += a[i] * b[i];
            s += sk;
        }
        return s;
    }
    double norm(const vector<double>& a) const { return std::sqrt(dot(a, a)); }
    static void axpy(double al, const vector<double>& x, vector<double>& y) {
        for (size_t i = 0; i < y.size(); ++i) y[i] += al * x[i];
    }
    static void scale(double al, vector<double>& x) {
        for (double& v : x) v *= al;
    }
    static double rho_est(double r0, double r, long it) {
        if (it > 1 && r0 > 0.0 && r > 0.0) return std::pow(r / r0, 1.0 / (double)(it - 1));
        return NAN;
    }

    // ----- in-group solve (sn_one_speed.jl:81-196) --------------------------------------
    void one_speed(int g, vector<double>& z_phi_bnd /* state; Φ part holds the group flux */,
                   const vector<double>& Qlout, oracle_stats& st) {
        vector<double> Ql(nphi);
        long passes = 0;
        bool any = false;
        for (double v : Qlout) if (v != 0.0) { any = true; break; }
        // the group's surface sources count too (sn_one_speed.jl:116: `sources`)
        const size_t npf[6] = {(size_t)Ny * Nz, (size_t)Ny * Nz, (size_t)Nx * Nz, (size_t)Nx * Nz,
                               (size_t)Nx * Ny, (size_t)Nx * Ny};
        for (int f = 0; f < 2 * ndims && !any; ++f) {
            if (!P.ssrc[f]) continue;
            const double* q = P.ssrc[f] + (size_t)g * P.Np_source * npf[f];
            for (size_t i = 0; i < (size_t)P.Np_source * npf[f]; ++i)
                if (q[i] != 0.0) { any = true; break; }
        }
        if (P.csd && !any)  // the incoming energy flux is a source too
            for (double v : psiE_in) if (v != 0.0) { any = true; break; }
        if (!any) {  // sn_one_speed.jl:116-120
            if (P.csd) psiE_out.assign((size_t)A.Nd * Nc * NmE, 0.0);
            std::fill(z_phi_bnd.begin(), z_phi_bnd.begin() + nphi, 0.0);
            st.niter = 0; st.resid = 0.0; st.rho = NAN; st.converged = 1; st.passes = 0;
            return;
        }
        vector<double>& work = z_phi_bnd;
        auto T = [&](vector<double>& out, const vector<double>& zin) {
            work = zin; pass(g, false, work, Qlout, Ql); out = work; ++passes;
        };
        auto matvec = [&](vector<double>& out, const vector<double>& zin) {
            work = zin; pass(g, true, work, Qlout, Ql); out = work; ++passes;
            scale(-1.0, out); axpy(1.0, zin, out);
        };
        string acc = P.accel;
        long niter = 0; double resid = 1.0, rho = NAN; bool conv = false;
        vector<double> z;
        if (acc == "none" || acc == "livolant") {
            z = work;
            livolant(z, T, acc == "none" ? -1 : 3, niter, resid, conv, rho);
        } else if (acc == "anderson") {
            z = work;
            anderson(z, T, niter, resid, conv, rho);
        } else if (acc == "gmres" || acc == "bicgstab") {
            std::fill(work.begin(), work.end(), 0.0);
            pass(g, false, work, Qlout, Ql); ++passes;
            vector<double> c = work;
            z.assign(S, 0.0);
            if (acc == "gmres") gmres(z, matvec, c, niter, resid, conv, rho);
            else bicgstab(z, matvec, c, niter, resid, conv, rho);
        } else {
            fail("Unknown acceleration method: " + acc + ".");
        }
        work = z;
        pass(g, false, work, Qlout, Ql); ++passes;  // reconstruction (sn_one_speed.jl:186)
        st.niter = niter; st.resid = resid; st.rho = rho; st.converged = conv; st.passes = (int32_t)passes;
        if (P.verbose)
            std::printf(">>>Group %d has %sconverged ( eps = %.4E , N = %ld , rho = %.4f )\n", g + 1,
                        conv ? "" : "not ", resid, niter, rho);
    }

    // tools/gmres.jl:35-133
    template <class MV>
    void gmres(vector<double>& z, MV matvec, const vector<double>& c, long& iter, double& resid, bool& conv,
               double& rho) {
        int m = std::max(1, (int)P.restart);
        long maxit = P.maxit;
        double tol = P.tol;
        double bnorm = norm(c);
        if (bnorm == 0.0) {
            std::fill(z.begin(), z.end(), 0.0);
            iter = 0; resid = 0.0; conv = true; rho = NAN;
            return;
        }
        vector<vector<double>> V(m + 1, vector<double>(S, 0.0));
        vector<double> w(S, 0.0), r(S, 0.0);
        vector<double> H((size_t)(m + 1) * m, 0.0);  // H[i + (m+1)*j]
        vector<double> cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0);
        auto h = [&](int i, int j) -> double& { return H[i + (size_t)(m + 1) * j]; };
        iter = 0; resid = 1.0; double resid0 = NAN; conv = false;
        while (iter < maxit && !conv) {
            matvec(r, z); ++iter;
            scale(-1.0, r); axpy(1.0, c, r);
            double beta = norm(r);
            resid = beta / bnorm;
            if (std::isnan(resid0)) resid0 = resid;
            if (resid < tol) { conv = true; break; }
            V[0] = r; scale(1.0 / beta, V[0]);
            std::fill(gv.begin(), gv.end(), 0.0);
            gv[0] = beta;
            int k = 0;
            for (int j = 0; j < m; ++j) {
                if (iter >= maxit) break;
                matvec(w, V[j]); ++iter;
                for (int i = 0; i <= j; ++i) {
                    h(i, j) = dot(w, V[i]);
                    axpy(-h(i, j), V[i], w);
                }
                double hjp1 = norm(w);
                h(j + 1, j) = hjp1;
                if (hjp1 > 0.0) { V[j + 1] = w; scale(1.0 / hjp1, V[j + 1]); }
                for (int i = 0; i < j; ++i) {
                    double t = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                    h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                    h(i, j) = t;
                }
                double den = std::hypot(h(j, j), h(j + 1, j));
                if (den == 0.0) { cs[j] = 1.0; sn[j] = 0.0; }
                else { cs[j] = h(j, j) / den; sn[j] = h(j + 1, j) / den; }
                h(j, j) = cs[j] * h(j, j) + sn[j] * h(j + 1, j);
                h(j + 1, j) = 0.0;
                gv[j + 1] = -sn[j] * gv[j];
                gv[j] = cs[j] * gv[j];
                k = j + 1;
                resid = std::fabs(gv[j + 1]) / bnorm;
                if (resid < tol) { conv = true; break; }
                if (hjp1 == 0.0) break;
            }
            vector<double> y(k, 0.0);
            for (int i = k - 1; i >= 0; --i) {
                double s = gv[i];
                for (int l = i + 1; l < k; ++l) s -= h(i, l) * y[l];
                y[i] = h(i, i) != 0.0 ? s / h(i, i) : 0.0;
            }
            for (int i = 0; i < k; ++i) axpy(y[i], V[i], z);
            if (conv) break;
        }
        rho = rho_est(resid0, resid, iter);
    }

    // tools/anderson.jl:38-140 (beta = 1)
    template <class FP>
    void anderson(vector<double>& z, FP fixedpoint, long& iter, double& resid, bool& conv, double& rho) {
        int m = std::max(1, (int)P.depth);
        double beta = 1.0;
        vector<double> g(S), f(S), f_old(S, 0.0), g_old(S, 0.0);
        vector<vector<double>> dF, dG;
        iter = 0; resid = 1.0; double resid0 = NAN; conv = false;
        for (long it = 1; it <= P.maxit; ++it) {
            fixedpoint(g, z); ++iter;
            f = g; axpy(-1.0, z, f);
            double fnorm = norm(f);
            double sc = std::max(std::max(norm(z), norm(g)), 1e-16);
            resid = fnorm / sc;
            if (std::isnan(resid0)) resid0 = resid;
            if (resid < P.tol) { conv = true; break; }
            if (it == 1) {
                axpy(beta, f, z);
            } else {
                vector<double> a = f; axpy(-1.0, f_old, a);
                vector<double> b = g; axpy(-1.0, g_old, b);
                dF.push_back(a); dG.push_back(b);
                if ((int)dF.size() > m) { dF.erase(dF.begin()); dG.erase(dG.begin()); }
                int mk = (int)dF.size();
                vector<double> Nm((size_t)mk * mk), bb(mk);
                for (int i = 0; i < mk; ++i) {
                    bb[i] = dot(dF[i], f);
                    for (int j = 0; j < mk; ++j) Nm[i + (size_t)mk * j] = dot(dF[i], dF[j]);
                }
                double tr = 0.0;
                for (int i = 0; i < mk; ++i) tr += Nm[i + (size_t)mk * i];
                double lam = 1e-12 * (tr / mk + 1e-300);
                for (int i = 0; i < mk; ++i) Nm[i + (size_t)mk * i] += lam;
                vector<double> gam = anderson_solve(Nm, bb, mk);
                axpy(beta, f, z);
                for (int i = 0; i < mk; ++i) {
                    axpy(-gam[i], dG[i], z);
                    axpy(-(beta - 1.0) * gam[i], dF[i], z);
                }
            }
            f_old = f; g_old = g;
        }
        rho = rho_est(resid0, resid, iter);
    }

    // tools/anderson.jl:109-140
    static vector<double> anderson_solve(vector<double> M, vector<double> x, int n) {
        for (int k = 0; k < n; ++k) {
            int p = k;
            double mx = std::fabs(M[k + (size_t)n * k]);
            for (int i = k + 1; i < n; ++i)
                if (std::fabs(M[i + (size_t)n * k]) > mx) { mx = std::fabs(M[i + (size_t)n * k]); p = i; }
            if (p != k) {
                for (int j = k; j < n; ++j) std::swap(M[k + (size_t)n * j], M[p + (size_t)n * j]);
                std::swap(x[k], x[p]);
            }
            double piv = M[k + (size_t)n * k];
            if (piv == 0.0) continue;
            for (int i = k + 1; i < n; ++i) {
                double fct = M[i + (size_t)n * k] / piv;
                for (int j = k; j < n; ++j) M[i + (size_t)n * j] -= fct * M[k + (size_t)n * j];
                x[i] -= fct * x[k];
            }
        }
        for (int k = n - 1; k >= 0; --k) {
            double s = x[k];
            for (int j = k + 1; j < n; ++j) s -= M[k + (size_t)n * j] * x[j];
            x[k] = M[k + (size_t)n * k] != 0.0 ? s / M[k + (size_t)n * k] : 0.0;
        }
        return x;
    }

    // tools/livolant.jl:69-81 and :115-155 (period < 0 means no extrapolation: "none")
    template <class FP>
    void livolant(vector<double>& z, FP fixedpoint, int period, long& iter, double& resid, bool& conv,
                  double& rho) {
        vector<double> g(S), f(S), z1(S, 0.0), z0(S, 0.0);
        int nhist = 0;
        iter = 0; resid = 1.0; double resid0 = NAN; conv = false;
        for (long it = 1; it <= P.maxit; ++it) {
            fixedpoint(g, z); ++iter;
            f = g; axpy(-1.0, z, f);
            double fnorm = norm(f);
            double sc = std::max(norm(g), 1e-16);
            resid = fnorm / sc;
            if (std::isnan(resid0)) resid0 = resid;
            if (resid < P.tol) { z = g; conv = true; break; }
            if (period > 0 && it % period == 0 && nhist >= 2) {
                size_t off[5] = {0, nphi, nphi + nbx, nphi + nbx + nby, S};
                double se = 0.0, sd = 0.0;
                for (int k = 0; k < 4; ++k) {
                    double a = 0.0, b = 0.0;
                    for (size_t i = off[k]; i < off[k + 1]; ++i) {
                        double e0 = z1[i] - z0[i];
                        double e1 = g[i] - z1[i];
                        double de = e1 - e0;
                        a += e0 * de;
                        b += de * de;
                    }
                    se += a; sd += b;
                }
                double mu = 1.0;
                if (sd != 0.0) { mu = -se / sd; if (mu <= 0) mu = 1.0; }
                z = g; scale(mu, z); axpy(1.0 - mu, z1, z);
            } else {
                z = g;
            }
            z0 = z1; z1 = g; ++nhist;
        }
        rho = rho_est(resid0, resid, iter);
    }

    // tools/bicgstab.jl:30-101
    template <class MV>
    void bicgstab(vector<double>& z, MV matvec, const vector<double>& c, long& iter, double& resid, bool& conv,
                  double& rho) {
        double bnorm = norm(c);
        if (bnorm == 0.0) { std::fill(z.begin(), z.end(), 0.0); iter = 0; resid = 0; conv = true; rho = NAN; return; }
        vector<double> r(S), r0(S), p(S, 0.0), v(S, 0.0), s(S), t(S);
        matvec(r, z);
        scale(-1.0, r); axpy(1.0, c, r);
        iter = 1;
        resid = norm(r) / bnorm;
        double resid0 = resid;
        conv = false;
        if (resid < P.tol) { conv = true; rho = NAN; return; }
        r0 = r;
        double rho_old = 1.0, al = 1.0, om = 1.0, tiny = 1e-300;
        while (iter < P.maxit) {
            double rn = dot(r0, r);
            if (std::fabs(rn) < tiny) break;
            double be = (rn / rho_old) * (al / om);
            axpy(-om, v, p); scale(be, p); axpy(1.0, r, p);
            matvec(v, p); ++iter;
            double den = dot(r0, v);
            if (std::fabs(den) < tiny) break;
            al = rn / den;
            s = r; axpy(-al, v, s);
            if (norm(s) / bnorm < P.tol) {
                axpy(al, p, z);
                resid = norm(s) / bnorm;
                conv = true;
                break;
            }
            matvec(t, s); ++iter;
            double tt = dot(t, t);
            if (tt < tiny) break;
            om = dot(t, s) / tt;
            axpy(al, p, z); axpy(om, s, z);
            r = s; axpy(-om, t, r);
            resid = norm(r) / bnorm;
            if (resid < P.tol) { conv = true; break; }
            if (std::fabs(om) < tiny) break;
            rho_old = rn;
        }
        rho = rho_est(resid0, resid, iter);
    }

    // sn_flux.jl:150-232 -- one Gauss-Seidel pass over groups (upscatter sees zero flux).
    void solve(double* phi_out, oracle_stats* stats, double* cutoff_out = nullptr) {
        const int Np = A.Np;
        vector<double> phi_all((size_t)G * nphi, 0.0);  // [g][p + Np*c]
        for (int g = 0; g < G; ++g) {
            vector<double> Qlout(nphi, 0.0);
            // scattering_source.jl:59-68 (all gi != gf, in order; none for solver types 4 and 5,
            // sn_flux.jl:176)
            for (int gi = 0; gi < G; ++gi) {
                if (gi == g || solver == 4 || solver == 5) continue;
                const double* pg = &phi_all[(size_t)gi * nphi];
                for (int c = 0; c < Ncol; ++c) {
                    int m = matc[c / K];
                    for (int p = 0; p < Np; ++p)
                        Qlout[p + (size_t)Np * c] += sig_s(m, gi, g, (int)A.B.pl[p]) * pg[p + (size_t)Np * c];
                }
            }
            // sn_flux.jl:179 -- volume source enters moment p=1 and spatial moment 1 only (Source.jl:135)
            for (int c = 0; c < Nc; ++c) Qlout[(size_t)Np * K * c] += P.qvol[g + (size_t)G * c];
            vector<double> z(S, 0.0);
            std::copy(phi_all.begin() + (size_t)g * nphi, phi_all.begin() + (size_t)(g + 1) * nphi, z.begin());
            if (P.csd) {  // sn_flux.jl:166,183: zero for the first group, then rescaled by ΔE_g/ΔE_{g-1}
                if (g == 0) psiE_in.assign((size_t)A.Nd * Nc * NmE, 0.0);
                else
                    for (size_t i = 0; i < psiE_in.size(); ++i) psiE_in[i] = psiE_out[i] * P.dE[g] / P.dE[g - 1];
            }
            oracle_stats st{};
            one_speed(g, z, Qlout, st);
            std::copy(z.begin(), z.begin() + nphi, phi_all.begin() + (size_t)g * nphi);
            if (stats) stats[g] = st;
        }
        // sn_flux.jl:213-218 -- flux at the cutoff energy: the energy-face angular flux leaving the
        // last group (not rescaled), projected on the moments with Dn, accumulated over n outermost.
        // Layout 𝚽cutoff[p, is, ix, iy, iz] with is over the cell moments; the energy face's NmE
        // spatial moments fill is = 1..NmE (sn_flux.jl:216), the others stay zero.
        if (cutoff_out) {
            std::fill(cutoff_out, cutoff_out + (size_t)Np * Ncol, 0.0);
            if (P.csd)
                for (int n = 0; n < A.Nd; ++n)
                    for (int c = 0; c < Nc; ++c)
                        for (int is = 0; is < NmE; ++is)
                            for (int p = 0; p < Np; ++p)
                                cutoff_out[p + (size_t)Np * (is + (size_t)K * c)] +=
                                    A.B.D[p + (size_t)Np * n] * psiE_out[is + (size_t)NmE * (c + (size_t)Nc * n)];
        }
        // Radiant layout 𝚽l[g, p, is, ix, iy, iz]
        for (int c = 0; c < Ncol; ++c)
            for (int p = 0; p < Np; ++p)
                for (int g = 0; g < G; ++g)
                    phi_out[g + (size_t)G * (p + (size_t)Np * c)] = phi_all[(size_t)g * nphi + p + (size_t)Np * c];
    }
};

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

// Quadrature: returns Nd (or a negative status); arrays must hold `cap` entries.
int oracle_quadrature(int N, const char* type, int Ndims, int Qdims, double* mu, double* eta, double* xi,
                      double* w, int cap) {
    try {
        vector<double> a, b, c, d;
        quadrature(N, type, Ndims, Qdims, a, b, c, d);
        int Nd = (int)d.size();
        if (Nd > cap) fail("capacity too small");
        for (int i = 0; i < Nd; ++i) { mu[i] = a[i]; eta[i] = b[i]; xi[i] = c[i]; w[i] = d[i]; }
        return Nd;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Volume basis: M is Nd x Np column-major, D is Np x Nd column-major; returns Np.
int oracle_basis(const double* mu, const double* eta, const double* xi, const double* w, int Nd, int L,
                 const char* type, int Qdims, double* M, double* D, int64_t* pl, int64_t* pm, int capNp) {
    try {
        vector<double> a(mu, mu + Nd), b(eta, eta + Nd), c(xi, xi + Nd), d(w, w + Nd);
        Basis B = angular_polynomial_basis(a, b, c, d, L, type, Qdims);
        if (B.Np > capNp) fail("capacity too small");
        std::copy(B.M.begin(), B.M.end(), M);
        std::copy(B.D.begin(), B.D.end(), D);
        for (int i = 0; i < B.Np; ++i) { pl[i] = B.pl[i]; pm[i] = B.pm[i]; }
        return B.Np;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Half-range basis of one face (0=X-,1=X+,2=Y-,3=Y+,4=Z-,5=Z+): returns Np_surf; *Ndp receives
// the number of directions entering the face.  M is Ndp x Nps, D is Nps x Ndp.
int oracle_surface_basis(const double* mu, const double* eta, const double* xi, const double* w, int Nd, int L,
                         const char* type, int Qdims, int face, double* M, double* D, int64_t* n_to_np,
                         int* Ndp, int cap) {
    try {
        vector<double> a(mu, mu + Nd), b(eta, eta + Nd), c(xi, xi + Nd), d(w, w + Nd);
        SurfBasis S = surface_basis(a, b, c, d, L, type, Qdims, face);
        if ((int)S.M.size() > cap || (int)S.D.size() > cap) fail("capacity too small");
        std::copy(S.M.begin(), S.M.end(), M);
        std::copy(S.D.begin(), S.D.end(), D);
        for (int n = 0; n < Nd; ++n) n_to_np[n] = S.n_to_np[n];
        *Ndp = S.Ndp;
        return S.Np;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Harmonics (for pinning tests): R has (L+1)^2 entries, R[l*l + l + m].
int oracle_real_sh(int L, double mu, double phi, int half_range, double* R) {
    try {
        Harm H = half_range ? real_half_range_sh_up_to_L(L, mu, phi) : real_sh_up_to_L(L, mu, phi);
        for (int l = 0; l <= L; ++l)
            for (int m = -l; m <= l; ++m) R[l * l + l + m] = H[l][l + m];
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

int oracle_legendre(int L, double x, int half_range, double* P) {
    try {
        vector<double> v = half_range ? half_range_legendre_up_to_L(L, x) : legendre_up_to_L(L, x);
        std::copy(v.begin(), v.end(), P);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Fokker-Planck operator of a problem (types 2 and 4): ℳ [Np x Np] column-major and λ₀;
// returns Np.
// The Voronoi neighbours of every node of a set on the unit sphere (voronoi_sphere): nbr receives
// the neighbour lists one after another, cnt[n] their lengths; returns the total (-1 on error or
// when it exceeds cap).
int oracle_voronoi_neighbours(int Nd, const double* x, const double* y, const double* z, int* cnt, int* nbr, int cap) {
    try {
        const auto nb = voronoi_neighbours(vector<double>(x, x + Nd), vector<double>(y, y + Nd), vector<double>(z, z + Nd));
        int tot = 0;
        for (int n = 0; n < Nd; ++n) {
            cnt[n] = (int)nb[n].size();
            for (int m : nb[n]) {
                if (tot >= cap) fail("capacity too small");
                nbr[tot++] = m;
            }
        }
        return tot;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// ℳ_EM of a problem with a magnetic field: [G][Np x Np]; returns Np.
int oracle_em_matrix(const oracle_problem* p, double* Mem, int cap) {
    try {
        Solver s(*p);
        if (s.Mem.empty()) fail("the problem has no electromagnetic field");
        if ((int)s.Mem.size() > cap) fail("capacity too small");
        std::copy(s.Mem.begin(), s.Mem.end(), Mem);
        return s.A.Np;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

double oracle_ferrer(int l, int m, double x) { return ferrer(l, m, x); }

// constant_linear(x, y): out = (Qx, Qy, Tx, Ty); returns the subdomain (0..11) or -1 on error.
int oracle_constant_linear(double x, double y, double* out) {
    try {
        cl::constant_linear(x, y, out, out + 2);
        return cl::which(x, y);
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

int oracle_fokker_planck(const oracle_problem* p, double* Mfp, double* lam0, int capNp) {
    try {
        Solver s(*p);
        if (s.Mfp.empty()) fail("the problem's solver type has no Fokker-Planck term");
        if (s.A.Np > capNp) fail("capacity too small");
        std::copy(s.Mfp.begin(), s.Mfp.end(), Mfp);
        *lam0 = s.lam0;
        return s.A.Np;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One 1D cell of scheme (DD, DG, DG-, DG+) and order K: psi[K] receives the spatial moments of the
// angular flux, *fx enters as the incoming face flux and leaves as the outgoing one.
int oracle_cell_1d(const char* scheme, int K, double mu, double St, double dx, const double* Qn, double* fx,
                   double* psi) {
    try {
        vector<double> om = scheme_weights_1d(scheme, K);
        flux_1d_bte(mu, St, dx, Qn, *fx, K, om.data(), psi, std::string(scheme) == "AWD");
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One 2D cell of closures (DD or DG per axis) and orders (Ox, Oy), fully coupled or not: psi receives
// the spatial moments, fx[Oy] and fy[Ox] enter as the incoming face moments and leave as the outgoing.
int oracle_cell_2d(const char* sx, const char* sy, int Ox, int Oy, int partial, double mu, double eta, double St,
                   double dx, double dy, const double* Qn, double* fx, double* fy, double* psi) {
    try {
        vector<double> wx = scheme_weights_1d(sx, Ox), wy = scheme_weights_1d(sy, Oy);
        const bool fc = partial == 0;
        Cell2 cl{Ox, Oy, fc ? Ox * Oy : Ox + Oy - 1, fc};
        const bool ad = std::string(sx) == "AWD" || std::string(sy) == "AWD";
        flux_2d_bte(mu, eta, St, dx, dy, Qn, fx, fy, cl, wx.data(), wy.data(), psi, ad);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One 3D cell of closures (DD or DG per axis) and orders: psi receives the spatial moments; fx, fy,
// fz enter as the incoming face moments (over (y, z), (x, z), (x, y)) and leave as the outgoing.
int oracle_cell_3d(const char* sx, const char* sy, const char* sz, int Ox, int Oy, int Oz, int partial,
                   const double* om, const double* h, double St, const double* Qn, double* fx, double* fy, double* fz,
                   double* psi) {
    try {
        vector<double> wx = scheme_weights_1d(sx, Ox), wy = scheme_weights_1d(sy, Oy), wz = scheme_weights_1d(sz, Oz);
        const double* w[3] = {wx.data(), wy.data(), wz.data()};
        Cell3 cl(Ox, Oy, Oz, partial == 0);
        const bool ad = std::string(sx) == "AWD" || std::string(sy) == "AWD" || std::string(sz) == "AWD";
        flux_3d_bte(om, h, St, Qn, fx, fy, fz, cl, w, psi, ad);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One slowing-down cell (flux_bfp) of the given closures and orders: axes E, x, y, z; ndims = 1
// applies the 1D tangential guard.  fe/fx/fy/fz: incoming face moments, overwritten with the
// outgoing ones; Sm, Sp: S⁻/ΔE and S⁺/ΔE (the in-group moments follow from them).
int oracle_cell_bfp(const char* se, const char* sx, const char* sy, const char* sz, int Oe, int Ox, int Oy, int Oz,
                    int partial, int ndims, const double* om, const double* h, double St, double Sm, double Sp,
                    const double* Qn, double* fe, double* fx, double* fy, double* fz, double* psi) {
    try {
        if (Oe < 1 || Oe > 2) fail("energy closures of order above 2 are not supported");
        double Sv[2];
        stopping_moments(Sm, Sp, Oe, Sv);
        for (const char* sc : {se, sx, sy, sz})
            if (string(sc) != "DD" && string(sc) != "DG") fail("Closure relation not implemented yet.");
        vector<double> we = scheme_weights_1d(se, Oe), wx = scheme_weights_1d(sx, Ox), wy = scheme_weights_1d(sy, Oy),
                       wz = scheme_weights_1d(sz, Oz);
        const double* w[4] = {we.data(), wx.data(), wy.data(), wz.data()};
        vector<double> W((size_t)Oe * Oe * Oe);
        for (int a = 0; a < Oe; ++a)
            for (int b = 0; b < Oe; ++b)
                for (int c = 0; c < Oe; ++c) W[a + Oe * (b + Oe * c)] = legendre_triple(a, b, c);
        Cell4 cl(Oe, Ox, Oy, Oz, partial == 0);
        double* f[4] = {fe, fx, fy, fz};
        flux_bfp(om, h, St, Sm, Sp, Sv, W.data(), Qn, f, cl, w, ndims == 1 && std::fabs(om[0]) <= 1e-10, psi);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One slowing-down cell of order 1 on every axis with the adaptive closure (AWD0 over x[, y[, z]]
// and E, with 1/ΔE as the energy axis' h): as oracle_cell_bfp, and dE the group width.
int oracle_cell_bfp_awd(int ndims, const double* om, const double* h, double St, double Sm, double Sp, double dE,
                        const double* Qn, double* fe, double* fx, double* fy, double* fz, double* psi) {
    try {
        double Sv[1];
        stopping_moments(Sm, Sp, 1, Sv);
        vector<double> we = scheme_weights_1d("AWD", 1), wx = we, wy = we, wz = we;
        const double* w[4] = {we.data(), wx.data(), wy.data(), wz.data()};
        const double W[1] = {legendre_triple(0, 0, 0)};
        Cell4 cl(1, 1, 1, 1, true);
        double* f[4] = {fe, fx, fy, fz};
        flux_bfp(om, h, St, Sm, Sp, Sv, W, Qn, f, cl, w, ndims == 1 && std::fabs(om[0]) <= 1e-10, psi, true, 1.0 / dE);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One 1D slowing-down cell of orders (2, 2) in (x, E), full coupling, with the coupled linear AWD1:
// Qn[4] (moment 2 ix + iE), fe[2] (energy face over x moments), fx[2] (x face over E moments) in and out.
int oracle_cell_bfp1d_awd1(double mu, double dx, double St, double Sm, double Sp, double dE, const double* Qn,
                           double* fe, double* fx, double* psi) {
    try {
        double Sv[2];
        stopping_moments(Sm, Sp, 2, Sv);
        vector<double> W(8);
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) W[a + 2 * (b + 2 * c)] = legendre_triple(a, b, c);
        flux_bfp1d_awd1(mu, dx, St, Sm, Sp, Sv, W.data(), Qn, fe, fx, 1.0 / dE, psi);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// ½ ∫ P_a P_b P_c dx (legendre_triple).
double oracle_legendre_triple(int a, int b, int c) { return legendre_triple(a, b, c); }

// Sizes of a problem: Nd, Np, Np_surf and the Krylov state length S.
int oracle_problem_sizes(const oracle_problem* p, int* Nd, int* Np, int* Nps, int64_t* S) {
    try {
        Solver s(*p);
        *Nd = s.A.Nd; *Np = s.A.Np; *Nps = s.A.Nps; *S = (int64_t)s.S;
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Full solve: phi_out in Radiant's 𝚽l[G, Np, 1, Nx, Ny, Nz] layout; stats_out[G].
int oracle_solve(const oracle_problem* p, double* phi_out, oracle_stats* stats_out) {
    try {
        Solver s(*p);
        s.solve(phi_out, stats_out);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// Full solve that also returns 𝚽cutoff[Np, 1, Nx, Ny, Nz] (sn_flux.jl:151,213-218,229) in
// cutoff_out (zeros unless continuous slowing down is on).
int oracle_solve_cutoff(const oracle_problem* p, double* phi_out, double* cutoff_out, oracle_stats* stats_out) {
    try {
        Solver s(*p);
        s.solve(phi_out, stats_out, cutoff_out);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// One transport pass of group g (0-based) on a state vector z of length S (in place), with
// out-of-group source Qlout [Np*Nc] (p fastest).  homogeneous=1 gives A z, 0 gives T(z).
int oracle_pass(const oracle_problem* p, int g, int homogeneous, double* z, const double* Qlout) {
    try {
        Solver s(*p);
        vector<double> zz(z, z + s.S), q(Qlout, Qlout + s.nphi), Ql(s.nphi);
        s.pass(g, homogeneous != 0, zz, q, Ql);
        std::copy(zz.begin(), zz.end(), z);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// oracle_pass timed: the wall time of the pass alone (the set-up of the solver excluded) in *sec.
int oracle_pass_timed(const oracle_problem* p, int g, int homogeneous, double* z, const double* Qlout, double* sec) {
    try {
        Solver s(*p);
        vector<double> zz(z, z + s.S), q(Qlout, Qlout + s.nphi), Ql(s.nphi);
        const double t0 = omp_get_wtime();
        s.pass(g, homogeneous != 0, zz, q, Ql);
        *sec = omp_get_wtime() - t0;
        std::copy(zz.begin(), zz.end(), z);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

}  // extern "C"
