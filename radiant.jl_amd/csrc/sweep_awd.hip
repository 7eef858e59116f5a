// sweep_awd.hip -- sweeps with the adaptive weighted diamond closure (SN scheme "AWD";
// particle_transport/adaptive.jl, scheme_weights.jl:32-37), FP64, for gfx950.
//
// The closure weights of a cell follow from its incoming face fluxes and source moments (Alcouffe's
// adaptive weighted diamond), recomputed cell by cell in every pass.  The outgoing face flux is then
// no longer an affine map of the incoming one, so a sweep walks its cells in order:
// - 1D (sn_sweep_1D.jl:51-124 with flux_1D_BTE, sn_1D_bte.jl:27-64): one lane per direction walks
//   its row of cells; orders 1 (AWD0, adaptive.jl:36-54) and 2 (AWD1, adaptive.jl:56-73).  A
//   direction with |mu| <= 1e-10 does not stream: its face flux passes through unchanged and the
//   weights are not adapted (sn_1D_bte.jl:39).
// - slowing-down types, order 1 on every axis (flux_{1,2,3}D_BFP): below, with the energy axis.
// - 2D and 3D (sn_sweep_2D.jl:54-200 / sn_sweep_3D.jl:57-240 with flux_2D_BTE / flux_3D_BTE): one
//   workgroup per direction walks the diagonals kx + ky (+ kz) = d in sweep order, one cell per
//   thread, with the face fluxes of every pencil in LDS (3D grids whose faces exceed it: in the
//   direction's slice of hoface); order 1 (AWD0, adaptive.jl:127-158 and 257-297).
// AWD0 on n axes: the diamond estimate
//   phi_DD = (Q + 2 h_1 f_1 + ... + 2 h_n f_n) / (2 h_1 + ... + 2 h_n + St),  h_a = |Omega_a| / Delta_a,
//   u_a = (f_a - phi_DD) / phi_DD,  P_a = 1 when 3|u_a| <= 1, else 1/(3|u_a|), and 0 when negative or NaN,
// gives the weights w_a = (-P_a, 1 + P_a) of the order-1 cell
//   psi = (Q - sum_a h_a (w_a0 - 1) f_a) / (St + sum_a h_a w_a1),   f_a <- w_a0 f_a + w_a1 psi.
// AWD1 (1D, order 2): dQ = sqrt3 s (Q_1 - sqrt3 h s f) / (Q_0 + h f),  Qx = (3 + dQ) / (3 (3 - dQ)) for
// dQ > 3/2, -1/dQ for dQ < -1, else 1; the weights (0, 1, Qx) enter the 2 x 2 cell of sn_1D_bte.jl
//   S_ij = St d_ij - [i > j] C_i h s C_j (1 - (-1)^(i-j)) + C_i h s^i C_j s^j w_(j+1),
//   Q_j = Qn_j - C_j h (s^j w_0 - (-s)^j) f,  f <- w_0 f + sum_j C_j s^j w_(j+1) psi_j   (C_j = sqrt(2j+1)).
// Source and flux moments: psi[n][k + K * cell], in place (source in, flux out); incoming face fluxes
// from bin_x/bin_y/bin_z ([Nd][pos], zero when null), exit faces into bx/by/bz.
#include "sweep_awd.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cmath>
#include <stdexcept>
#include <string>

namespace rt {

#define RT_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                       " at " __FILE__ ":" + std::to_string(__LINE__));   \
    } while (0)

// The AWD0 weights of n axes: w[a][0] = -P_a, w[a][1] = 1 + P_a (adaptive.jl:36-54, 127-158, 257-297).
template <int N>
__device__ __forceinline__ void awd0(const double* h, const double* f, double Q, double St, double (*w)[2]) {
    double num = Q, den = 2.0 * h[0];
#pragma unroll
    for (int a = 0; a < N; ++a) num += 2.0 * h[a] * f[a];
#pragma unroll
    for (int a = 1; a < N; ++a) den += 2.0 * h[a];
    const double phi_dd = num / (den + St);
#pragma unroll
    for (int a = 0; a < N; ++a) {
        const double u = (f[a] - phi_dd) / phi_dd;
        double P = 3.0 * fabs(u) <= 1.0 ? 1.0 : 1.0 / (3.0 * fabs(u));
        if (P < 0.0 || isnan(P)) P = 0.0;
        w[a][0] = -P;
        w[a][1] = 1.0 + P;
    }
}

// ------------------------------------------------------------------------------------------- 1D
template <int K>
__global__ __launch_bounds__(64) void sweep_1d_awd_kernel(const PassDesc* __restrict__ pds) {
    const PassDesc& d = pds[0];
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= d.Nd) return;
    const int Nx = d.Nx, Nc = d.Nc;
    double mu = d.mu[n];
    const bool fwd = mu >= 0.0;
    const bool tang = fabs(mu) <= 1e-10;
    const double sx = (tang || mu > 0.0) ? 1.0 : -1.0;
    if (tang) mu = 0.0;
    double om0[K + 1], Cc[K], sp[K], sn[K];  // base weights, C_j, sx^j, (-sx)^j
#pragma unroll
    for (int i = 0; i <= K; ++i) om0[i] = d.ho_om[i];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        Cc[i] = sqrt(2.0 * i + 1.0);
        sp[i] = i == 0 ? 1.0 : sp[i - 1] * sx;
        sn[i] = i == 0 ? 1.0 : -sn[i - 1] * sx;
    }
    double* __restrict__ psi = (double*)d.psi + (size_t)n * Nc;
    double f = d.bin_x ? d.bin_x[n] : 0.0;
    for (int k = 0; k < Nx; ++k) {
        const int ix = fwd ? k : Nx - 1 - k;
        const double hx = fabs(mu) / d.dx[ix];
        const double St = d.sigt[d.mat[(size_t)K * ix]];
        double Qn[K], om[K + 1];
#pragma unroll
        for (int j = 0; j < K; ++j) Qn[j] = psi[j + (size_t)K * ix];
#pragma unroll
        for (int j = 0; j <= K; ++j) om[j] = om0[j];
        if (!tang) {
            if constexpr (K == 1) {
                double w[1][2];
                awd0<1>(&hx, &f, Qn[0], St, w);
                om[0] = w[0][0];
                om[1] = w[0][1];
            } else {
                const double r3 = sqrt(3.0);
                const double dQ = r3 * sx * (Qn[1] - r3 * hx * sx * f) / (Qn[0] + hx * f);
                double Qx;
                if (dQ > 1.5) Qx = (3.0 + dQ) / (3.0 * (3.0 - dQ));
                else if (dQ < -1.0) Qx = -1.0 / dQ;
                else Qx = 1.0;
                om[0] = 0.0;
                om[1] = 1.0;
                om[2] = Qx + 0.0;
            }
        }
        double S[K][K], q[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double e = i == j ? St : 0.0;
                if (i > j && ((i - j) & 1)) e -= Cc[i] * hx * sx * Cc[j] * 2.0;
                e += Cc[i] * hx * sp[i] * Cc[j] * sp[j] * om[j + 1];
                S[i][j] = e;
            }
            q[i] = Qn[i] - Cc[i] * hx * (sp[i] * om[0] - sn[i]) * f;
        }
        if constexpr (K == 2) {  // partial pivoting, then elimination
            if (fabs(S[1][0]) > fabs(S[0][0])) {
                for (int j = 0; j < 2; ++j) { const double t = S[0][j]; S[0][j] = S[1][j]; S[1][j] = t; }
                const double t = q[0]; q[0] = q[1]; q[1] = t;
            }
            const double fct = S[1][0] / S[0][0];
            S[1][1] -= fct * S[0][1];
            q[1] -= fct * q[0];
            q[1] = q[1] / S[1][1];
            q[0] = (q[0] - S[0][1] * q[1]) / S[0][0];
        } else {
            q[0] = q[0] / S[0][0];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) psi[j + (size_t)K * ix] = q[j];
        if (!tang) {
            double fo = om[0] * f;
#pragma unroll
            for (int j = 0; j < K; ++j) fo += Cc[j] * sp[j] * om[j + 1] * q[j];
            f = fo;
        }
    }
    if (d.bx) ((double*)d.bx)[n] = f;  // exit face flux (one position in 1D)
}

// ------------------------------------------------------------------------------------------- 2D
__global__ __launch_bounds__(256) void sweep_2d_awd_kernel(const PassDesc* __restrict__ pds) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny;
    double* FX = (double*)smem;  // [Ny]
    double* FY = FX + Ny;        // [Nx]
    const double mu = d.mu[n], eta = d.eta[n];
    const bool fxs = mu >= 0, fys = eta >= 0;
    double* __restrict__ psi = (double*)d.psi + (size_t)n * d.Nc;
    for (int i = tid; i < Ny; i += nt) FX[i] = d.bin_x ? d.bin_x[(size_t)n * Ny + i] : 0.0;
    for (int i = tid; i < Nx; i += nt) FY[i] = d.bin_y ? d.bin_y[(size_t)n * Nx + i] : 0.0;
    __syncthreads();
    for (int dg = 0; dg < Nx + Ny - 1; ++dg) {
        const int k0 = dg - (Ny - 1) > 0 ? dg - (Ny - 1) : 0, k1 = dg < Nx - 1 ? dg : Nx - 1;
        for (int kx = k0 + tid; kx <= k1; kx += nt) {
            const int ky = dg - kx;
            const int ix = fxs ? kx : Nx - 1 - kx, iy = fys ? ky : Ny - 1 - ky;
            const size_t c = (size_t)ix + (size_t)Nx * iy;
            const double h[2] = {fabs(mu) / d.dx[ix], fabs(eta) / d.dy[iy]};
            const double St = d.sigt[d.mat[c]];
            const double f[2] = {FX[iy], FY[ix]};
            const double Q = psi[c];
            double w[2][2];
            awd0<2>(h, f, Q, St, w);
            const double S = St + h[0] * w[0][1] + h[1] * w[1][1];
            const double q = Q - h[0] * (w[0][0] - 1.0) * f[0] - h[1] * (w[1][0] - 1.0) * f[1];
            const double v = q / S;
            psi[c] = v;
            FX[iy] = w[0][0] * f[0] + w[0][1] * v;
            FY[ix] = w[1][0] * f[1] + w[1][1] * v;
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny; i += nt) ((double*)d.bx)[(size_t)n * Ny + i] = FX[i];
    if (d.by)
        for (int i = tid; i < Nx; i += nt) ((double*)d.by)[(size_t)n * Nx + i] = FY[i];
}

// ------------------------------------------------------------------- 2D coupled linear (AWD1)
// Orders (2, 2) with full coupling (adaptive.jl:161-232 with constant_linear.jl): per cell the
// slopes of the source and inflows place the centroid of the cell's constant part, which sets
// (Qx, Qy, Tx, Ty) by subdomain of [-1/2, 1/2]^2; the weights are tensors w[n][m][k] (Radiant's
// ω[n+1, m+1, k+1]): those of DG with w[2][m][m] = Q and w[2][0][1] = s T / sqrt3 of the other axis'
// sign.  The cell's 4 x 4 system (moment jx + 2 jy) couples through w[jx+1][jy][iy], and an outgoing
// face moment m gathers sum_k w[jx+1][m][k] (sn_2D_bte.jl:35-116).  Faces [Ny][2] and [Nx][2] in LDS.
namespace awd1 {
struct cx { double re, im; };
__device__ __forceinline__ cx add(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cx mul(cx a, cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cx dvd(cx a, cx b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ cx cpow(cx z, double p) {  // principal value
    const double r = hypot(z.re, z.im), t = atan2(z.im, z.re);
    const double m = pow(r, p);
    return {m * cos(p * t), m * sin(p * t)};
}
__device__ double y_in(double x) {
    const double r = sqrt(36 * x * x - 60 * x + 9);
    return (-6 * x - 3 + r) / (36 * x - 30 - 6 * r);
}
__device__ double y_out(double x) { return (6 * x * x - 3 * x) / (6 * x + 1); }
// 0 A++ 1 A+- 2 A-+ 3 A-- 4 Bx+ 5 Bx- 6 By+ 7 By- 8 C++ 9 C+- 10 C-+ 11 C--
__device__ int which(double x, double y) {
    const double s6 = 1.0 / 6.0;
    if (x <= -s6) return y <= -s6 ? 11 : (y < s6 ? 5 : 10);
    if (x <= 0) {
        const double yo = y_out(-x), yi = y_in(-x);
        if (y <= yo) return 7;
        if (y < -yi) return 3;
        if (y <= yi) return 5;
        if (y < -yo) return 2;
        return 6;
    }
    if (x < s6) {
        const double yo = y_out(x), yi = y_in(x);
        if (y <= yo) return 7;
        if (y < -yi) return 1;
        if (y <= yi) return 4;
        if (y < -yo) return 0;
        return 6;
    }
    return y <= -s6 ? 9 : (y < s6 ? 4 : 8);
}
__device__ void dA(double x, double y, double& d1, double& d2) {
    if (fabs(x) < 1e-10 || fabs(y) < 1e-10) { d1 = 0; d2 = 0; return; }
    const double a = 2 * y / x, b = -6 * y / x + 3 - 6 * y, c = 9 * y / (2 * x) - 4.5 - 9 * x + 9 * y, d = 12 * x;
    const double D0 = b * b - 3 * a * c, D1 = 2 * b * b * b - 9 * a * b * c + 27 * a * a * d;
    const cx xi = {-0.5, -sqrt(3.0) / 2};
    const cx sq = cpow({D1 * D1 - 4 * D0 * D0 * D0, 0.0}, 0.5);
    const cx C = mul(cpow({(D1 + sq.re) / 2, sq.im / 2}, 1.0 / 3.0), xi);
    if (D0 == 0 && D1 == 0) d1 = -b / (3 * a);
    else {
        const cx t = add(add({b, 0.0}, C), dvd({D0, 0.0}, C));
        d1 = -t.re / (3 * a);
    }
    d2 = (2 * d1 - 3) * y / (2 * x) + 1.5;
}
__device__ double QAp(double i, double j) { return (6 - 2 * i * i * j) / (3 * i * i * i * j + 6 * (1 - i * i * j)); }
__device__ double QAm(double i, double j) { return (2 * i * i * j + 6 * (1 - i * j)) / (3 * i * i * i * j + 6 * (1 - i * i * j)); }
__device__ double QBp(double i, double j) {
    if (fabs(i - j) < 1e-5) return (1 + i) / ((1 - i) * (1 + 2 * i));
    return (1 - (i * i + j * j + i * j) / 3) / (1 - (i * i + j * j + i * j) + 0.5 * (i * i * i + j * j * j + i * i * j + i * j * j));
}
__device__ double QB0(double i, double j) {
    if (fabs(i - j) < 1e-5 && fabs(i - 1) < 1e-5) return 1.0;
    return 2.0 / 3.0 * (i + 2 * j - 3) / (i + j - 2);
}
__device__ double QBm(double i, double j) {
    if (fabs(i - j) < 1e-5) return 1 / (1 + 2 * i);
    return (1 - i - j + (i * i + j * j + i * j) / 3) / (1 - (i * i + j * j + i * j) + 0.5 * (i * i * i + j * j * j + i * i * j + i * j * j));
}
__device__ double TAden(double i, double j) { const double v = i * i * i * j - 2 * i * i * j + 2; return 5 * v * v; }
__device__ double TAp(double i, double j) { return -i * i * j * (i * i * i * j * j + 12 * i * j - 30 * i - 20 * j + 40) / TAden(i, j); }
__device__ double TAm(double i, double j) {
    return -i * j * (i * i * i * i * j * j - 8 * i * i * i * j * j + (10 * j * j - 12 * j + 30) * i * i + (40 * j - 80) * i - 40 * j + 60) /
           TAden(i, j);
}
__device__ double TBden(double i, double j) {
    const double v = i * i * i + (j - 2) * i * i + (j * j - 2 * j) * i + j * j * j - 2 * j * j + 2;
    return 5 * v * v;
}
__device__ double TBp(double i, double j) {
    if (fabs(i + j - 2) < 1e-5) return 0.0;
    return -(i - j) * ((i * i * i * i + 4 * i * i * i * j + 10 * i * i * j * j + 4 * i * j * j * j + j * j * j * j) -
                       6 * (3 * i * i + 4 * i * j + 3 * j * j) + 20 * (i + j)) / TBden(i, j);
}
__device__ double TBm(double i, double j) {
    if (fabs(i - j) < 1e-5) return 0.0;
    return (i - j) * ((i * i * i * i + 4 * i * i * i * j + 10 * i * i * j * j + 4 * i * j * j * j + j * j * j * j) -
                      8 * (i * i * i + 4 * i * i * j + 4 * i * j * j + j * j * j) + 4 * (7 * i * i + 16 * i * j + 7 * j * j) -
                      40 * (i + j) + 20) / TBden(i, j);
}
__device__ double TB0(double i, double j) {
    if (fabs(i - j) < 1e-5) return 0.0;
    return -(i - j) * (i * i + 6 * i * j + 3 * j * j - 8 * i - 12 * j + 10) / (5 * (i + j - 2) * (i + j - 2));
}
__device__ double TCp(double i, double j) {
    if (fabs(i - j) < 1e-5) return -(i * i + 6 * i + 3) / (5 * (1 + i) * (1 + i));
    if (fabs(i - 1) < 1e-5) return -INFINITY;
    return -(i * i + 6 * i + 3) * (j - 1) / (5 * (1 + i) * (1 + i) * (i - 1));
}
__device__ double TCm(double i, double j) { return -(i - 1) * (j - 1) / (5 * (1 + i) * (1 + i)); }
__device__ void constant_linear(double x, double y, double* Q, double* T) {
    const int sub = which(x, y);
    double d1, d2;
    switch (sub) {
        case 0: dA(x, y, d1, d2); break;
        case 1: dA(x, -y, d1, d2); break;
        case 2: dA(-x, y, d1, d2); break;
        case 3: dA(-x, -y, d1, d2); break;
        case 4: d1 = (12 * y * y + (6 - 12 * x) * y + 2 * x) / (12 * y * y + 1); d2 = (12 * y * y - (6 - 12 * x) * y + 2 * x) / (12 * y * y + 1); break;
        case 5: d1 = (12 * y * y + (6 + 12 * x) * y - 2 * x) / (12 * y * y + 1); d2 = (12 * y * y - (6 + 12 * x) * y - 2 * x) / (12 * y * y + 1); break;
        case 6: d1 = (12 * x * x + (6 - 12 * y) * x + 2 * y) / (12 * x * x + 1); d2 = (12 * x * x - (6 - 12 * y) * x + 2 * y) / (12 * x * x + 1); break;
        case 7: d1 = (12 * x * x + (6 + 12 * y) * x - 2 * y) / (12 * x * x + 1); d2 = (12 * x * x - (6 + 12 * y) * x - 2 * y) / (12 * x * x + 1); break;
        case 8: d1 = 3 * x - 0.5; d2 = 3 * y - 0.5; break;
        case 9: d1 = 3 * x - 0.5; d2 = -3 * y - 0.5; break;
        case 10: d1 = -3 * x - 0.5; d2 = 3 * y - 0.5; break;
        default: d1 = -3 * x - 0.5; d2 = -3 * y - 0.5; break;
    }
    if (d1 > 1 && d1 < 1.0000001) d1 = 1;
    if (d1 < 0 && d1 > -0.0000001) d1 = 0;
    if (d2 > 1 && d2 < 1.0000001) d2 = 1;
    if (d2 < 0 && d2 > -0.0000001) d2 = 0;
    switch (sub) {
        case 0: Q[0] = QAp(d1, d2); Q[1] = QAp(d2, d1); T[0] = TAp(d1, d2); T[1] = TAp(d2, d1); break;
        case 1: Q[0] = QAp(d1, d2); Q[1] = QAm(d2, d1); T[0] = -TAp(d1, d2); T[1] = -TAm(d2, d1); break;
        case 2: Q[0] = QAm(d1, d2); Q[1] = QAp(d2, d1); T[0] = -TAm(d1, d2); T[1] = -TAp(d2, d1); break;
        case 3: Q[0] = QAm(d1, d2); Q[1] = QAm(d2, d1); T[0] = TAm(d1, d2); T[1] = TAm(d2, d1); break;
        case 4: Q[0] = QBp(d1, d2); Q[1] = QB0(d1, d2); T[0] = TBp(d1, d2); T[1] = TB0(d1, d2); break;
        case 5: Q[0] = QBm(d1, d2); Q[1] = QB0(d1, d2); T[0] = TBm(d1, d2); T[1] = -TB0(d1, d2); break;
        case 6: Q[0] = QB0(d1, d2); Q[1] = QBp(d1, d2); T[0] = TB0(d1, d2); T[1] = TBp(d1, d2); break;
        case 7: Q[0] = QB0(d1, d2); Q[1] = QBm(d1, d2); T[0] = -TB0(d1, d2); T[1] = TBm(d1, d2); break;
        case 8: Q[0] = 2 * (2 + d1) / (3 * (1 - d1 * d1)); Q[1] = 2 * (2 + d2) / (3 * (1 - d2 * d2)); T[0] = TCp(d1, d2); T[1] = TCp(d2, d1); break;
        case 9: Q[0] = 2 * (2 + d1) / (3 * (1 - d1 * d1)); Q[1] = 2 / (3 * (1 + d2)); T[0] = -TCp(d1, d2); T[1] = -TCm(d2, d1); break;
        case 10: Q[0] = 2 / (3 * (1 + d1)); Q[1] = 2 * (2 + d2) / (3 * (1 - d2 * d2)); T[0] = -TCm(d1, d2); T[1] = -TCp(d2, d1); break;
        default: Q[0] = 2 / (3 * (1 + d1)); Q[1] = 2 / (3 * (1 + d2)); T[0] = TCm(d1, d2); T[1] = TCm(d2, d1); break;
    }
}
// the centroid coordinate of one axis from its slope estimate (adaptive.jl:172-209)
__device__ double centroid(double p0, double dQ) {
    if (p0 != 0) dQ = dQ / p0;
    else if (fabs(dQ) < 1e-8) dQ = 0;
    else dQ = dQ > 0 ? INFINITY : -INFINITY;
    if (isnan(dQ)) dQ = 0;
    if (dQ > 2.999) dQ = 2.999;
    if (dQ < -2.999) dQ = -2.999;
    const double xmax = 0.49, sg = dQ > 0 ? 1.0 : (dQ < 0 ? -1.0 : 0.0), dQt = dQ;
    double x = 0.0;
    dQ += 1.5 * sg;
    if (dQ > 1.5) x = dQ > 3 ? xmax : fmin(xmax, (2 * dQ - 3 + sqrt(12 * dQ * dQ - 27)) / (4 * (3 * dQ)));
    dQ = dQt + sg;
    if (dQ < -1) x = fmax(-xmax, (1 + dQ) / 4);
    return fmin(fmax(x, -xmax), xmax);
}
}  // namespace awd1

__global__ __launch_bounds__(256) void sweep_2d_awd1_kernel(const PassDesc* __restrict__ pds) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny;
    double* FX = (double*)smem;  // [Ny][2]
    double* FY = FX + 2 * Ny;    // [Nx][2]
    const double mu = d.mu[n], eta = d.eta[n];
    const bool fxs = mu >= 0, fys = eta >= 0;
    const double sx = mu > 0 ? 1.0 : (mu < 0 ? -1.0 : 0.0), sy = eta > 0 ? 1.0 : (eta < 0 ? -1.0 : 0.0);
    const double r3 = sqrt(3.0);
    const double Cc[2] = {1.0, r3}, px[2] = {1.0, sx}, nx[2] = {1.0, -sx}, py[2] = {1.0, sy}, ny[2] = {1.0, -sy};
    double* __restrict__ psi = (double*)d.psi + (size_t)n * d.Nc;
    for (int i = tid; i < 2 * Ny; i += nt) FX[i] = d.bin_x ? d.bin_x[(size_t)n * Ny * 2 + i] : 0.0;
    for (int i = tid; i < 2 * Nx; i += nt) FY[i] = d.bin_y ? d.bin_y[(size_t)n * Nx * 2 + i] : 0.0;
    __syncthreads();
    for (int dg = 0; dg < Nx + Ny - 1; ++dg) {
        const int k0 = dg - (Ny - 1) > 0 ? dg - (Ny - 1) : 0, k1 = dg < Nx - 1 ? dg : Nx - 1;
        for (int kx = k0 + tid; kx <= k1; kx += nt) {
            const int ky = dg - kx;
            const int ix = fxs ? kx : Nx - 1 - kx, iy = fys ? ky : Ny - 1 - ky;
            const size_t c = (size_t)ix + (size_t)Nx * iy;
            const double hx = fabs(mu) / d.dx[ix], hy = fabs(eta) / d.dy[iy];
            const double St = d.sigt[d.mat[c * 4]];
            double* fx = FX + 2 * (size_t)iy;
            double* fy = FY + 2 * (size_t)ix;
            double Qn[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Qn[k] = psi[(size_t)k + 4 * c];
            // weights
            const double p0 = Qn[0] + hx * fx[0] + hy * fy[0];
            const double xc = awd1::centroid(p0, sx * r3 * (Qn[1] - r3 * hx * sx * fx[0] + hy * fy[1]));
            const double yc = awd1::centroid(p0, sy * r3 * (Qn[2] + hx * fx[1] - r3 * hy * sy * fy[0]));
            double Qw[2], Tw[2];
            awd1::constant_linear(xc, yc, Qw, Tw);
            double wx[3][2][2], wy[3][2][2];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 2; ++e) wx[a][b][e] = wy[a][b][e] = (b == e && a > 0) ? 1.0 : 0.0;
            wx[2][0][0] = Qw[0]; wx[2][1][1] = Qw[0]; wx[2][0][1] = sy * Tw[0] / r3;
            wy[2][0][0] = Qw[1]; wy[2][1][1] = Qw[1]; wy[2][0][1] = sx * Tw[1] / r3;
            // the 4 x 4 system, moment jx + 2 jy
            double S[4][4], q[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) S[i][j] = i == j ? St : 0.0;
#pragma unroll
            for (int ax = 0; ax < 2; ++ax)
#pragma unroll
                for (int ay = 0; ay < 2; ++ay) {
                    const int i = ax + 2 * ay;
#pragma unroll
                    for (int bx = 0; bx < 2; ++bx)
#pragma unroll
                        for (int by = 0; by < 2; ++by) {
                            const int j = bx + 2 * by;
                            if (ay == by && ax > bx && ((ax - bx) & 1)) S[i][j] -= 2.0 * Cc[ax] * hx * sx * Cc[bx];
                            S[i][j] += Cc[ax] * hx * px[ax] * Cc[bx] * px[bx] * wx[bx + 1][by][ay];
                            if (ax == bx && ay > by && ((ay - by) & 1)) S[i][j] -= 2.0 * Cc[ay] * hy * sy * Cc[by];
                            S[i][j] += Cc[ay] * hy * py[ay] * Cc[by] * py[by] * wy[by + 1][bx][ax];
                        }
                    q[i] = Qn[i] - Cc[ax] * hx * (px[ax] * wx[0][ay][ay] - nx[ax]) * fx[ay] -
                           Cc[ay] * hy * (py[ay] * wy[0][ax][ax] - ny[ay]) * fy[ax];
                }
#pragma unroll
            for (int p = 0; p < 4; ++p) {
#pragma unroll
                for (int i = p + 1; i < 4; ++i) {
                    const bool sw = fabs(S[i][p]) > fabs(S[p][p]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double t = S[p][j];
                        S[p][j] = sw ? S[i][j] : t;
                        S[i][j] = sw ? t : S[i][j];
                    }
                    const double tq = q[p];
                    q[p] = sw ? q[i] : tq;
                    q[i] = sw ? tq : q[i];
                }
#pragma unroll
                for (int i = p + 1; i < 4; ++i) {
                    const double f = S[i][p] / S[p][p];
#pragma unroll
                    for (int j = p + 1; j < 4; ++j) S[i][j] -= f * S[p][j];
                    q[i] -= f * q[p];
                }
            }
#pragma unroll
            for (int i = 3; i >= 0; --i) {
                double v = q[i];
#pragma unroll
                for (int j = i + 1; j < 4; ++j) v -= S[i][j] * q[j];
                q[i] = v / S[i][i];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) psi[(size_t)k + 4 * c] = q[k];
            double ox[2] = {wx[0][0][0] * fx[0], wx[0][1][1] * fx[1]}, oy[2] = {wy[0][0][0] * fy[0], wy[0][1][1] * fy[1]};
#pragma unroll
            for (int ax = 0; ax < 2; ++ax)
#pragma unroll
                for (int ay = 0; ay < 2; ++ay) {
                    const double v = q[ax + 2 * ay];
                    ox[ay] += Cc[ax] * px[ax] * (wx[ax + 1][ay][0] + wx[ax + 1][ay][1]) * v;
                    oy[ax] += Cc[ay] * py[ay] * (wy[ay + 1][ax][0] + wy[ay + 1][ax][1]) * v;
                }
            fx[0] = ox[0]; fx[1] = ox[1];
            fy[0] = oy[0]; fy[1] = oy[1];
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < 2 * Ny; i += nt) ((double*)d.bx)[(size_t)n * Ny * 2 + i] = FX[i];
    if (d.by)
        for (int i = tid; i < 2 * Nx; i += nt) ((double*)d.by)[(size_t)n * Nx * 2 + i] = FY[i];
}

// ------------------------------------------------ 1D slowing-down cell of orders (2, 2) in (x, E)
// flux_1D_BFP (sn_1D_bfp.jl:39-125) with the coupled linear AWD1 over the axes (x, E): h = (hx, 1/ΔE),
// signs (sx, -1), the x face and the energy face as inflows, the cell's moments (2 ix + iE, energy
// fastest) handed to the closure as they are.  One lane per direction walks its row; the x face
// carries the 2 energy moments ([Nd][2] in bin_x / bx), the energy face of each cell the 2 x moments
// ([Nd][cells][2] in psiE_in / psiE_out).  The in-group stopping-power term is 2 sqrt3 Sv[jE] in the
// rows iE = 1 with Sv = ((S- + S+)/2, (S- - S+)/(2 sqrt3)).  A tangential direction keeps DG and its x face.
__global__ __launch_bounds__(64) void sweep_1d_bfp_awd1_kernel(const PassDesc* __restrict__ pds, double inv_dE) {
    const PassDesc& d = pds[0];
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= d.Nd) return;
    const int Nx = d.Nx;
    const size_t Ncell = (size_t)Nx;
    const double mu = d.mu[n];
    const bool tang = fabs(mu) <= 1e-10;
    const bool fwd = mu >= 0.0;
    const double sx = (tang || mu > 0.0) ? 1.0 : -1.0;
    const double r3 = sqrt(3.0);
    const double Cc[2] = {1.0, r3}, px[2] = {1.0, sx}, nx[2] = {1.0, -sx}, pe[2] = {1.0, -1.0};
    double* __restrict__ psi = (double*)d.psi + (size_t)n * d.Nc;
    const double* Ein = d.psiE_in ? d.psiE_in + (size_t)n * Ncell * 2 : nullptr;
    double* Eout = d.psiE_out ? d.psiE_out + (size_t)n * Ncell * 2 : nullptr;
    double fx[2] = {d.bin_x ? d.bin_x[(size_t)n * 2] : 0.0, d.bin_x ? d.bin_x[(size_t)n * 2 + 1] : 0.0};
    for (int k = 0; k < Nx; ++k) {
        const int ix = fwd ? k : Nx - 1 - k;
        const double hx = tang ? 0.0 : fabs(mu) / d.dx[ix];
        const int m = d.mat[(size_t)4 * ix];
        const double St = d.sigt[m], Sm = d.csdS[2 * m], Sp = d.csdS[2 * m + 1];
        const double Sv[2] = {(Sm + Sp) / 2.0, (Sm - Sp) / (2.0 * r3)};
        double Qn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) Qn[j] = psi[(size_t)j + 4 * ix];
        const double fE[2] = {Ein ? Ein[(size_t)ix * 2] : 0.0, Ein ? Ein[(size_t)ix * 2 + 1] : 0.0};
        double wx[3][2][2], wE[3][2][2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) wx[a][b][e] = wE[a][b][e] = (b == e && a > 0) ? 1.0 : 0.0;
        if (!tang) {
            const double p0 = Qn[0] + hx * fx[0] + inv_dE * fE[0];
            const double xc = awd1::centroid(p0, sx * r3 * (Qn[1] - r3 * hx * sx * fx[0] + inv_dE * fE[1]));
            const double ec = awd1::centroid(p0, -r3 * (Qn[2] + hx * fx[1] + r3 * inv_dE * fE[0]));
            double Qw[2], Tw[2];
            awd1::constant_linear(xc, ec, Qw, Tw);
            wx[2][0][0] = Qw[0]; wx[2][1][1] = Qw[0]; wx[2][0][1] = -Tw[0] / r3;
            wE[2][0][0] = Qw[1]; wE[2][1][1] = Qw[1]; wE[2][0][1] = sx * Tw[1] / r3;
        }
        double S[4][4], q[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) S[i][j] = i == j ? St : 0.0;
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
            for (int ae = 0; ae < 2; ++ae) {
                const int i = 2 * ax + ae;
#pragma unroll
                for (int bx = 0; bx < 2; ++bx)
#pragma unroll
                    for (int be = 0; be < 2; ++be) {
                        const int j = 2 * bx + be;
                        if (ae == be && ax > bx && ((ax - bx) & 1)) S[i][j] -= 2.0 * Cc[ax] * hx * sx * Cc[bx];
                        S[i][j] += Cc[ax] * hx * px[ax] * Cc[bx] * px[bx] * wx[bx + 1][be][ae];
                        if (ax == bx && ae == 1) S[i][j] += 2.0 * Cc[1] * Sv[be];
                        S[i][j] += Cc[ae] * Sp * pe[ae] * Cc[be] * pe[be] * wE[be + 1][bx][ax];
                    }
                q[i] = Qn[i] - Cc[ax] * hx * (px[ax] * wx[0][ae][ae] - nx[ax]) * fx[ae] -
                       Cc[ae] * (pe[ae] * Sp * wE[0][ax][ax] - Sm) * fE[ax];
            }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
#pragma unroll
            for (int i = p + 1; i < 4; ++i) {
                const bool sw = fabs(S[i][p]) > fabs(S[p][p]);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double t = S[p][j];
                    S[p][j] = sw ? S[i][j] : t;
                    S[i][j] = sw ? t : S[i][j];
                }
                const double tq = q[p];
                q[p] = sw ? q[i] : tq;
                q[i] = sw ? tq : q[i];
            }
#pragma unroll
            for (int i = p + 1; i < 4; ++i) {
                const double f = S[i][p] / S[p][p];
#pragma unroll
                for (int j = p + 1; j < 4; ++j) S[i][j] -= f * S[p][j];
                q[i] -= f * q[p];
            }
        }
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            double v = q[i];
#pragma unroll
            for (int j = i + 1; j < 4; ++j) v -= S[i][j] * q[j];
            q[i] = v / S[i][i];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) psi[(size_t)j + 4 * ix] = q[j];
        double ox[2] = {wx[0][0][0] * fx[0], wx[0][1][1] * fx[1]}, oE[2] = {wE[0][0][0] * fE[0], wE[0][1][1] * fE[1]};
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
            for (int ae = 0; ae < 2; ++ae) {
                const double v = q[2 * ax + ae];
                ox[ae] += Cc[ax] * px[ax] * (wx[ax + 1][ae][0] + wx[ax + 1][ae][1]) * v;
                oE[ax] += Cc[ae] * pe[ae] * (wE[ae + 1][ax][0] + wE[ae + 1][ax][1]) * v;
            }
        if (!tang) { fx[0] = ox[0]; fx[1] = ox[1]; }
        if (Eout) { Eout[(size_t)ix * 2] = oE[0]; Eout[(size_t)ix * 2 + 1] = oE[1]; }
    }
    if (d.bx) { ((double*)d.bx)[(size_t)n * 2] = fx[0]; ((double*)d.bx)[(size_t)n * 2 + 1] = fx[1]; }
}

// ------------------------------------------------------------------------------------------- 3D
__global__ __launch_bounds__(256) void sweep_3d_awd_kernel(const PassDesc* __restrict__ pds) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
    // face fluxes in LDS, or in this direction's slice of hoface (the waves of the workgroup share
    // one CU, so __syncthreads orders their device-memory writes and reads too)
    const size_t nface = (size_t)Ny * Nz + (size_t)Nx * Nz + (size_t)Nx * Ny;
    double* FX = d.hoface ? (double*)d.hoface + (size_t)n * nface : (double*)smem;  // [Ny*Nz]
    double* FY = FX + (size_t)Ny * Nz;                                             // [Nx*Nz]
    double* FZ = FY + (size_t)Nx * Nz;                                             // [Nx*Ny]
    const double om[3] = {d.mu[n], d.eta[n], d.xi[n]};
    const bool fw[3] = {om[0] >= 0, om[1] >= 0, om[2] >= 0};
    double* __restrict__ psi = (double*)d.psi + (size_t)n * d.Nc;
    for (int i = tid; i < Ny * Nz; i += nt) FX[i] = d.bin_x ? d.bin_x[(size_t)n * Ny * Nz + i] : 0.0;
    for (int i = tid; i < Nx * Nz; i += nt) FY[i] = d.bin_y ? d.bin_y[(size_t)n * Nx * Nz + i] : 0.0;
    for (int i = tid; i < Nx * Ny; i += nt) FZ[i] = d.bin_z ? d.bin_z[(size_t)n * Nx * Ny + i] : 0.0;
    __syncthreads();
    for (int dg = 0; dg < Nx + Ny + Nz - 2; ++dg) {
        for (int i = tid; i < Nx * Ny; i += nt) {
            const int kx = i % Nx, ky = i / Nx, kz = dg - kx - ky;
            if (kz < 0 || kz >= Nz) continue;
            const int ix = fw[0] ? kx : Nx - 1 - kx, iy = fw[1] ? ky : Ny - 1 - ky, iz = fw[2] ? kz : Nz - 1 - kz;
            const size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
            const double h[3] = {fabs(om[0]) / d.dx[ix], fabs(om[1]) / d.dy[iy], fabs(om[2]) / d.dz[iz]};
            const double St = d.sigt[d.mat[c]];
            double* fx = FX + ((size_t)iy + (size_t)Ny * iz);
            double* fy = FY + ((size_t)ix + (size_t)Nx * iz);
            double* fz = FZ + ((size_t)ix + (size_t)Nx * iy);
            const double f[3] = {*fx, *fy, *fz};
            const double Q = psi[c];
            double w[3][2];
            awd0<3>(h, f, Q, St, w);
            const double S = St + h[0] * w[0][1] + h[1] * w[1][1] + h[2] * w[2][1];
            const double q = Q - h[0] * (w[0][0] - 1.0) * f[0] - h[1] * (w[1][0] - 1.0) * f[1] -
                             h[2] * (w[2][0] - 1.0) * f[2];
            const double v = q / S;
            psi[c] = v;
            *fx = w[0][0] * f[0] + w[0][1] * v;
            *fy = w[1][0] * f[1] + w[1][1] * v;
            *fz = w[2][0] * f[2] + w[2][1] * v;
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * Nz; i += nt) ((double*)d.bx)[(size_t)n * Ny * Nz + i] = FX[i];
    if (d.by)
        for (int i = tid; i < Nx * Nz; i += nt) ((double*)d.by)[(size_t)n * Nx * Nz + i] = FY[i];
    if (d.bz)
        for (int i = tid; i < Nx * Ny; i += nt) ((double*)d.bz)[(size_t)n * Nx * Ny + i] = FZ[i];
}

// ------------------------------------------------------------------------ slowing-down cells
// flux_{1,2,3}D_BFP (sn_1D_bfp.jl:39-125, sn_2D_bfp.jl, sn_3D_bfp.jl) of order 1 on every axis with
// the adaptive closure: AWD0 over the axes (x[, y[, z]], E) with h_E = 1/ΔE and the energy face as the
// energy axis' inflow (as the reference calls adaptive), then the order-1 slowing-down cell
//   S = St + sum_a h_a w_a1 + S+ wE1,   Q = Qn - sum_a h_a (w_a0 - 1) f_a - (S+ wE0 - S-) f_E,
//   psi = Q / S,  f_a <- w_a0 f_a + w_a1 psi,  f_E <- wE0 f_E + wE1 psi
// (S-, S+: the stopping powers at the group's upper and lower energy over ΔE).  A tangential 1D
// direction keeps the base weights (sn_1D_bfp.jl:52) and its x face.  The walk, the faces (in LDS or
// hoface, [Ny*Nz] + [Nx*Nz] (2D, 3D) + [Nx*Ny] (3D)) and the energy faces ([Nd][cells], psiE_in /
// psiE_out, absent in homogeneous passes) are those of the general slowing-down walker.
__global__ __launch_bounds__(256) void sweep_bfp_awd_kernel(const PassDesc* __restrict__ pds, double inv_dE) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int nd = d.ndims, Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
    const size_t Ncell = (size_t)Nx * Ny * Nz;
    const size_t nface = (size_t)Ny * Nz + (nd >= 2 ? (size_t)Nx * Nz : 0) + (nd >= 3 ? (size_t)Nx * Ny : 0);
    double* FX = d.hoface ? (double*)d.hoface + (size_t)n * nface : (double*)smem;  // [Ny*Nz]
    double* FY = FX + (size_t)Ny * Nz;                                             // [Nx*Nz] (2D, 3D)
    double* FZ = FY + (nd >= 2 ? (size_t)Nx * Nz : 0);                             // [Nx*Ny] (3D)
    const double om[3] = {d.mu[n], nd >= 2 ? d.eta[n] : 0.0, nd >= 3 ? d.xi[n] : 0.0};
    const bool tang = nd == 1 && fabs(om[0]) <= 1e-10;
    const bool fw[3] = {om[0] >= 0, om[1] >= 0, om[2] >= 0};
    double* __restrict__ psi = (double*)d.psi + (size_t)n * d.Nc;
    const double* Ein = d.psiE_in ? d.psiE_in + (size_t)n * Ncell : nullptr;
    double* Eout = d.psiE_out ? d.psiE_out + (size_t)n * Ncell : nullptr;
    for (int i = tid; i < Ny * Nz; i += nt) FX[i] = d.bin_x ? d.bin_x[(size_t)n * Ny * Nz + i] : 0.0;
    if (nd >= 2)
        for (int i = tid; i < Nx * Nz; i += nt) FY[i] = d.bin_y ? d.bin_y[(size_t)n * Nx * Nz + i] : 0.0;
    if (nd >= 3)
        for (int i = tid; i < Nx * Ny; i += nt) FZ[i] = d.bin_z ? d.bin_z[(size_t)n * Nx * Ny + i] : 0.0;
    __syncthreads();
    for (int dg = 0; dg < Nx + Ny + Nz - 2; ++dg) {
        for (int i = tid; i < Nx * Ny; i += nt) {
            const int kx = i % Nx, ky = i / Nx, kz = dg - kx - ky;
            if (kz < 0 || kz >= Nz) continue;
            const int ix = fw[0] ? kx : Nx - 1 - kx, iy = fw[1] ? ky : Ny - 1 - ky, iz = fw[2] ? kz : Nz - 1 - kz;
            const size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
            const double h[3] = {tang ? 0.0 : fabs(om[0]) / d.dx[ix], nd >= 2 ? fabs(om[1]) / d.dy[iy] : 0.0,
                                 nd >= 3 ? fabs(om[2]) / d.dz[iz] : 0.0};
            const int m = d.mat[c];
            const double St = d.sigt[m], Sm = d.csdS[2 * m], Sp = d.csdS[2 * m + 1];
            double* fx = FX + ((size_t)iy + (size_t)Ny * iz);
            double* fy = nd >= 2 ? FY + ((size_t)ix + (size_t)Nx * iz) : nullptr;
            double* fz = nd >= 3 ? FZ + ((size_t)ix + (size_t)Nx * iy) : nullptr;
            const double f[3] = {*fx, fy ? *fy : 0.0, fz ? *fz : 0.0};
            const double fe = Ein ? Ein[c] : 0.0;
            const double Q = psi[c];
            double w[4][2];  // x, y, z, E
            if (!tang) {
                const double h4[4] = {h[0], h[1], h[2], inv_dE}, f4[4] = {f[0], f[1], f[2], fe};
                awd0<4>(h4, f4, Q, St, w);
            } else {
                const double* base[4] = {d.ho_om, d.ho_omy, d.ho_omz, d.ho_ome};
#pragma unroll
                for (int a = 0; a < 4; ++a) { w[a][0] = base[a][0]; w[a][1] = base[a][1]; }
            }
            const double S = St + h[0] * w[0][1] + h[1] * w[1][1] + h[2] * w[2][1] + Sp * w[3][1];
            const double q = Q - h[0] * (w[0][0] - 1.0) * f[0] - h[1] * (w[1][0] - 1.0) * f[1] -
                             h[2] * (w[2][0] - 1.0) * f[2] - (Sp * w[3][0] - Sm) * fe;
            const double v = q / S;
            psi[c] = v;
            if (!tang) *fx = w[0][0] * f[0] + w[0][1] * v;
            if (fy) *fy = w[1][0] * f[1] + w[1][1] * v;
            if (fz) *fz = w[2][0] * f[2] + w[2][1] * v;
            if (Eout) Eout[c] = w[3][0] * fe + w[3][1] * v;
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * Nz; i += nt) ((double*)d.bx)[(size_t)n * Ny * Nz + i] = FX[i];
    if (nd >= 2 && d.by)
        for (int i = tid; i < Nx * Nz; i += nt) ((double*)d.by)[(size_t)n * Nx * Nz + i] = FY[i];
    if (nd >= 3 && d.bz)
        for (int i = tid; i < Nx * Ny; i += nt) ((double*)d.bz)[(size_t)n * Nx * Ny + i] = FZ[i];
}

template <typename F>
static void awd_launch(F fn, dim3 grid, dim3 block, size_t lds, hipStream_t st, const PassDesc* d_pd) {
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    if (a) hipExtLaunchKernelGGL(fn, grid, block, (uint32_t)lds, st, a, b, 0u, d_pd);
    else hipLaunchKernelGGL(fn, grid, block, (uint32_t)lds, st, d_pd);
}

bool sweep_awd_supported(int ndims, int Kx, int Ky, int Kz) {
    if (ndims == 1) return Kx == 1 || Kx == 2;
    if (ndims == 2) return (Kx == 1 && Ky == 1) || (Kx == 2 && Ky == 2);  // (2, 2): full coupling
    return ndims == 3 && Kx == 1 && Ky == 1 && Kz == 1;
}

void launch_sweep_awd(const PassDesc* d_pd, int Nd, int ndims, int K, size_t lds, hipStream_t st) {
    if (Nd <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("adaptive-closure face fluxes exceed the LDS");
    if (ndims == 1) {
        const dim3 grid((unsigned)((Nd + 63) / 64));
        if (K == 1) awd_launch(sweep_1d_awd_kernel<1>, grid, dim3(64), 0, st, d_pd);
        else if (K == 2) awd_launch(sweep_1d_awd_kernel<2>, grid, dim3(64), 0, st, d_pd);
        else throw std::runtime_error("1D adaptive closure of order other than 1 or 2");
    } else {
        if (K != 1 && !(ndims == 2 && K == 4))
            throw std::runtime_error("adaptive closure: order 1 in 2D/3D or (2, 2) with full coupling in 2D");
        auto fn = ndims == 3 ? sweep_3d_awd_kernel : (K == 4 ? sweep_2d_awd1_kernel : sweep_2d_awd_kernel);
        const int slot = ndims == 3 ? 2 : (K == 4 ? 1 : 0);
        static bool attr[3] = {false, false, false};
        if (!attr[slot]) {
            RT_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr[slot] = true;
        }
        awd_launch(fn, dim3(Nd), dim3(256), lds, st, d_pd);
    }
    RT_HIP(hipGetLastError());
}

void launch_sweep_1d_bfp_awd1(const PassDesc* d_pd, int Nd, double inv_dE, hipStream_t st) {
    if (Nd <= 0) return;
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    const dim3 grid((unsigned)((Nd + 63) / 64));
    if (a) hipExtLaunchKernelGGL(sweep_1d_bfp_awd1_kernel, grid, dim3(64), 0, st, a, b, 0u, d_pd, inv_dE);
    else hipLaunchKernelGGL(sweep_1d_bfp_awd1_kernel, grid, dim3(64), 0, st, d_pd, inv_dE);
    RT_HIP(hipGetLastError());
}

void launch_sweep_bfp_awd(const PassDesc* d_pd, int Nd, double inv_dE, size_t lds, hipStream_t st) {
    if (Nd <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("adaptive slowing-down face fluxes exceed the LDS");
    static bool attr = false;
    if (!attr) {
        RT_HIP(hipFuncSetAttribute((const void*)sweep_bfp_awd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   160 * 1024));
        attr = true;
    }
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    if (a) hipExtLaunchKernelGGL(sweep_bfp_awd_kernel, dim3(Nd), dim3(256), (uint32_t)lds, st, a, b, 0u, d_pd, inv_dE);
    else hipLaunchKernelGGL(sweep_bfp_awd_kernel, dim3(Nd), dim3(256), (uint32_t)lds, st, d_pd, inv_dE);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
