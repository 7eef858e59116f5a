// sweep_bfp.hip -- slowing-down sweeps (SN solver types BFP, BCSD, FP, CSD, BFP-EF) with a general
// closure in space and energy, for gfx950.
//
// flux_1D_BFP / flux_2D_BFP / flux_3D_BFP (sn_1D_bfp.jl:44-125, sn_2D_bfp.jl:43-154,
// sn_3D_bfp.jl:51-191; non-adaptive) with the sweeps of sn_sweep_{1,2,3}D.jl: one workgroup per
// direction walks the diagonal planes kx + ky + kz = d of the cell grid in sweep order; the cells
// of a plane are independent, one per thread.  The face moments of every x, y and z pencil live in
// LDS; the energy face of each cell comes in from psiE_in (the previous group's outflow, rescaled;
// absent in homogeneous passes) and leaves into psiE_out.  1D and 2D problems run the same walk
// with one position on the axes they lack (no streaming there).
//
// A cell holds the Legendre moments over the axes (E, x, y, z) with orders (KE, KX, KY, KZ):
// with full coupling moment (e, a, b, c) sits at e + KE*(a + KX*(b + KY*c)), energy fastest;
// without it only one index is non-zero and sits at that index plus the (order - 1) of the axes
// before it.  A face carries the moments over the other three axes by the same rule.  With
// C_k = sqrt(2k+1), 0-based orders, h_a = |Ω_a|/Δ_a, s_a = sign(Ω_a), energy streaming towards
// lower energy (sign -1) with S⁺ (the stopping power at the group's lower energy, per unit group
// width) as its rate, the cell system is
//   S_ij = St δ_ij
//        + Σ_axes [other indices equal] ( -[u>v, u-v odd] 2 C_u h s C_v + C_u h s^u C_v s^v w_(v+1) )
//        + [spatial indices equal] ( [e_i = 1] 2 C_1 Sv_(e_j) + C_ei S⁺ (-1)^ei C_ej (-1)^ej wE_(ej+1) )
//   Q_j  = Qn_j - Σ_axes C_v h (s^v w_0 - (-s)^v) f_axis - C_e ((-1)^e S⁺ wE_0 - S⁻) f_E
// where Sv = ((S⁻+S⁺)/2, (S⁻-S⁺)/(2√3)) are the in-group stopping-power moments (sn_flux.jl:92-96):
// for energy orders up to 2 the in-group term Σ_{k<i,w} C_i C_j C_k C_w (1-(-1)^(i-k)) Sv_w 𝒲_jkw of
// the reference reduces to that closed form (only k = 0 occurs, and 𝒲_j0w = δ_jw / (2j+1)).  The
// outgoing faces follow the closure f <- w_0 f + Σ_v C_v s^v w_(v+1) ψ (energy: (-1)^e for s^v).
// In 1D a direction with |μ| <= 1e-10 does not stream and keeps its x face (sn_1D_bfp.jl:52-54).
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdexcept>
#include <string>

namespace rt {

#define RT_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                       " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

namespace {

template <int KE, int KX, int KY, int KZ, bool FC>
struct Ho4 {
    __host__ __device__ static constexpr int ord(int a) { return a == 0 ? KE : (a == 1 ? KX : (a == 2 ? KY : KZ)); }
    static constexpr int NM = FC ? KE * KX * KY * KZ : KE + KX + KY + KZ - 3;
    __host__ __device__ static constexpr int face(int skip) {
        int prod = 1, sum = 0;
        for (int k = 0; k < 4; ++k)
            if (k != skip) { prod *= ord(k); sum += ord(k) - 1; }
        return FC ? prod : 1 + sum;
    }
    __host__ __device__ static constexpr bool valid(int e, int a, int b, int c) {
        return FC || (e > 0) + (a > 0) + (b > 0) + (c > 0) <= 1;
    }
    // index of (e, a, b, c) over the axes other than skip (-1: all four)
    __host__ __device__ static constexpr int index(int e, int a, int b, int c, int skip) {
        const int m[4] = {e, a, b, c};
        int r = 0, stride = 1, off = 0;
        for (int k = 0; k < 4; ++k) {
            if (k == skip) continue;
            if (FC) {
                r += m[k] * stride;
                stride *= ord(k);
            } else {
                if (m[k] > 0) r = m[k] + off;
                off += ord(k) - 1;
            }
        }
        return r;
    }
};

template <typename T, int KE, int KX, int KY, int KZ, bool FC>
__global__ __launch_bounds__(256) void sweep_bfp_kernel(const PassDesc* __restrict__ pds) {
    using H = Ho4<KE, KX, KY, KZ, FC>;
    constexpr int NM = H::NM, FE = H::face(0), F0 = H::face(1), F1 = H::face(2), F2 = H::face(3);
    constexpr int KS = KX > KY ? (KX > KZ ? KX : KZ) : (KY > KZ ? KY : KZ);
    constexpr int KM = KS > KE ? KS : KE;
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int nd = d.ndims, Nx = d.Nx, Ny = d.Ny, Nz = d.Nz, Nc = d.Nc;
    const size_t Ncell = (size_t)Nx * Ny * Nz;
    // face moments in LDS, or in this direction's slice of hoface (the waves of the workgroup share
    // one CU, so __syncthreads orders their device-memory writes and reads as it does for LDS)
    const size_t nface = (size_t)Ny * Nz * F0 + (nd >= 2 ? (size_t)Nx * Nz * F1 : 0) + (nd >= 3 ? (size_t)Nx * Ny * F2 : 0);
    T* FX = d.hoface ? (T*)d.hoface + (size_t)n * nface : (T*)smem;  // [Ny*Nz][F0]
    T* FY = FX + (size_t)Ny * Nz * F0;                    // [Nx*Nz][F1] (2D, 3D)
    T* FZ = FY + (nd >= 2 ? (size_t)Nx * Nz * F1 : 0);    // [Nx*Ny][F2] (3D)
    const double om[3] = {d.mu[n], nd >= 2 ? d.eta[n] : 0.0, nd >= 3 ? d.xi[n] : 0.0};
    const bool tang = nd == 1 && fabs(om[0]) <= 1e-10;
    const int K3[3] = {KX, KY, KZ};
    T s[3], w[3][KM + 1], we[KM + 1], Cc[KM + 1], sp[3][KM], sn[3][KM], pe[KM];
    const double* ws[3] = {d.ho_om, d.ho_omy, d.ho_omz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a] = om[a] > 0 ? T(1) : (om[a] < 0 ? T(-1) : T(0));
        if (a == 0 && tang) s[a] = T(1);
#pragma unroll
        for (int k = 0; k <= KM; ++k) w[a][k] = k <= K3[a] ? (T)ws[a][k] : T(0);
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            sp[a][k] = k == 0 ? T(1) : sp[a][k - 1] * s[a];
            sn[a][k] = k == 0 ? T(1) : -sn[a][k - 1] * s[a];
        }
    }
#pragma unroll
    for (int k = 0; k <= KM; ++k) we[k] = k <= KE ? (T)d.ho_ome[k] : T(0);
#pragma unroll
    for (int k = 0; k <= KM; ++k) Cc[k] = (T)sqrt(2.0 * k + 1.0);
#pragma unroll
    for (int k = 0; k < KM; ++k) pe[k] = (k & 1) ? T(-1) : T(1);
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    const double* Ein = d.psiE_in ? d.psiE_in + (size_t)n * Ncell * FE : nullptr;
    double* Eout = d.psiE_out ? d.psiE_out + (size_t)n * Ncell * FE : nullptr;
    for (int i = tid; i < Ny * Nz * F0; i += nt) FX[i] = d.bin_x ? (T)d.bin_x[(size_t)n * Ny * Nz * F0 + i] : T(0);
    if (nd >= 2)
        for (int i = tid; i < Nx * Nz * F1; i += nt) FY[i] = d.bin_y ? (T)d.bin_y[(size_t)n * Nx * Nz * F1 + i] : T(0);
    if (nd >= 3)
        for (int i = tid; i < Nx * Ny * F2; i += nt) FZ[i] = d.bin_z ? (T)d.bin_z[(size_t)n * Nx * Ny * F2 + i] : T(0);
    __syncthreads();
    const bool fw[3] = {om[0] >= 0, om[1] >= 0, om[2] >= 0};
    const T sq3 = (T)1.7320508075688772;
    for (int dg = 0; dg < Nx + Ny + Nz - 2; ++dg) {
        for (int i = tid; i < Nx * Ny; i += nt) {
            const int kx = i % Nx, ky = i / Nx, kz = dg - kx - ky;
            if (kz < 0 || kz >= Nz) continue;
            const int ix = fw[0] ? kx : Nx - 1 - kx, iy = fw[1] ? ky : Ny - 1 - ky, iz = fw[2] ? kz : Nz - 1 - kz;
            const size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
            const T h[3] = {tang ? T(0) : (T)(fabs(om[0]) / d.dx[ix]), nd >= 2 ? (T)(fabs(om[1]) / d.dy[iy]) : T(0),
                            nd >= 3 ? (T)(fabs(om[2]) / d.dz[iz]) : T(0)};
            const int m = d.mat[c * NM];
            const T St = (T)d.sigt[m];
            const T Sm = (T)d.csdS[2 * m], Sp = (T)d.csdS[2 * m + 1];
            const T Sv[2] = {(Sm + Sp) / T(2), (Sm - Sp) / (T(2) * sq3)};
            T zero[F1 > F2 ? F1 : F2];  // the faces of the axes a 1D or 2D problem lacks
#pragma unroll
            for (int k = 0; k < (F1 > F2 ? F1 : F2); ++k) zero[k] = T(0);
            T* fx = FX + ((size_t)iy + (size_t)Ny * iz) * F0;
            T* fy = nd >= 2 ? FY + ((size_t)ix + (size_t)Nx * iz) * F1 : zero;
            T* fz = nd >= 3 ? FZ + ((size_t)ix + (size_t)Nx * iy) * F2 : zero;
            T fe[FE];
#pragma unroll
            for (int k = 0; k < FE; ++k) fe[k] = Ein ? (T)Ein[c * FE + k] : T(0);
            T S[NM][NM], q[NM];
#pragma unroll
            for (int i2 = 0; i2 < NM; ++i2)
#pragma unroll
                for (int j = 0; j < NM; ++j) S[i2][j] = i2 == j ? St : T(0);
#pragma unroll
            for (int e = 0; e < KE; ++e)
#pragma unroll
            for (int a = 0; a < KX; ++a)
#pragma unroll
            for (int b = 0; b < KY; ++b)
#pragma unroll
            for (int cz = 0; cz < KZ; ++cz) {
                if (!H::valid(e, a, b, cz)) continue;
                const int ii = H::index(e, a, b, cz, -1);
                const int ia[3] = {a, b, cz};
#pragma unroll
                for (int e2 = 0; e2 < KE; ++e2)
#pragma unroll
                for (int a2 = 0; a2 < KX; ++a2)
#pragma unroll
                for (int b2 = 0; b2 < KY; ++b2)
#pragma unroll
                for (int c2 = 0; c2 < KZ; ++c2) {
                    if (!H::valid(e2, a2, b2, c2)) continue;
                    const int jj = H::index(e2, a2, b2, c2, -1);
                    const int jb[3] = {a2, b2, c2};
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) {
                        const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                        if (e != e2 || ia[o1] != jb[o1] || ia[o2] != jb[o2]) continue;
                        const int u = ia[ax], v = jb[ax];
                        if (u > v && ((u - v) & 1)) S[ii][jj] -= T(2) * Cc[u] * h[ax] * s[ax] * Cc[v];
                        S[ii][jj] += Cc[u] * h[ax] * sp[ax][u] * Cc[v] * sp[ax][v] * w[ax][v + 1];
                    }
                    if (a == a2 && b == b2 && cz == c2) {
                        if (e == 1) S[ii][jj] += T(2) * Cc[1] * Sv[e2];
                        S[ii][jj] += Cc[e] * Sp * pe[e] * Cc[e2] * pe[e2] * we[e2 + 1];
                    }
                }
                T r = psi[(size_t)ii + NM * c];
                r -= Cc[a] * h[0] * (sp[0][a] * w[0][0] - sn[0][a]) * fx[H::index(e, a, b, cz, 1)];
                if (nd >= 2) r -= Cc[b] * h[1] * (sp[1][b] * w[1][0] - sn[1][b]) * fy[H::index(e, a, b, cz, 2)];
                if (nd >= 3) r -= Cc[cz] * h[2] * (sp[2][cz] * w[2][0] - sn[2][cz]) * fz[H::index(e, a, b, cz, 3)];
                r -= Cc[e] * (pe[e] * Sp * we[0] - Sm) * fe[H::index(e, a, b, cz, 0)];
                q[ii] = r;
            }
            // Gaussian elimination with partial pivoting (row swaps as selects)
#pragma unroll
            for (int p = 0; p < NM; ++p) {
#pragma unroll
                for (int i2 = p + 1; i2 < NM; ++i2) {
                    const bool sw = fabs(S[i2][p]) > fabs(S[p][p]);
#pragma unroll
                    for (int j = 0; j < NM; ++j) {
                        const T t = S[p][j];
                        S[p][j] = sw ? S[i2][j] : t;
                        S[i2][j] = sw ? t : S[i2][j];
                    }
                    const T tq = q[p];
                    q[p] = sw ? q[i2] : tq;
                    q[i2] = sw ? tq : q[i2];
                }
#pragma unroll
                for (int i2 = p + 1; i2 < NM; ++i2) {
                    const T f = S[i2][p] / S[p][p];
#pragma unroll
                    for (int j = p + 1; j < NM; ++j) S[i2][j] -= f * S[p][j];
                    q[i2] -= f * q[p];
                }
            }
#pragma unroll
            for (int i2 = NM - 1; i2 >= 0; --i2) {
                T v = q[i2];
#pragma unroll
                for (int j = i2 + 1; j < NM; ++j) v -= S[i2][j] * q[j];
                q[i2] = v / S[i2][i2];
            }
#pragma unroll
            for (int i2 = 0; i2 < NM; ++i2) psi[(size_t)i2 + NM * c] = q[i2];
            // outgoing faces
            T ox[F0], oy[F1], oz[F2], oe[FE];
#pragma unroll
            for (int k = 0; k < F0; ++k) ox[k] = w[0][0] * fx[k];
#pragma unroll
            for (int k = 0; k < F1; ++k) oy[k] = nd >= 2 ? w[1][0] * fy[k] : T(0);
#pragma unroll
            for (int k = 0; k < F2; ++k) oz[k] = nd >= 3 ? w[2][0] * fz[k] : T(0);
#pragma unroll
            for (int k = 0; k < FE; ++k) oe[k] = we[0] * fe[k];
#pragma unroll
            for (int e = 0; e < KE; ++e)
#pragma unroll
            for (int a = 0; a < KX; ++a)
#pragma unroll
            for (int b = 0; b < KY; ++b)
#pragma unroll
            for (int cz = 0; cz < KZ; ++cz) {
                if (!H::valid(e, a, b, cz)) continue;
                const T v = q[H::index(e, a, b, cz, -1)];
                ox[H::index(e, a, b, cz, 1)] += Cc[a] * sp[0][a] * w[0][a + 1] * v;
                oy[H::index(e, a, b, cz, 2)] += Cc[b] * sp[1][b] * w[1][b + 1] * v;
                oz[H::index(e, a, b, cz, 3)] += Cc[cz] * sp[2][cz] * w[2][cz + 1] * v;
                oe[H::index(e, a, b, cz, 0)] += Cc[e] * pe[e] * we[e + 1] * v;
            }
            if (!tang)
#pragma unroll
                for (int k = 0; k < F0; ++k) fx[k] = ox[k];
            if (nd >= 2)
#pragma unroll
                for (int k = 0; k < F1; ++k) fy[k] = oy[k];
            if (nd >= 3)
#pragma unroll
                for (int k = 0; k < F2; ++k) fz[k] = oz[k];
            if (Eout)
#pragma unroll
                for (int k = 0; k < FE; ++k) Eout[c * FE + k] = (double)oe[k];
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * Nz * F0; i += nt) ((T*)d.bx)[(size_t)n * Ny * Nz * F0 + i] = FX[i];
    if (nd >= 2 && d.by)
        for (int i = tid; i < Nx * Nz * F1; i += nt) ((T*)d.by)[(size_t)n * Nx * Nz * F1 + i] = FY[i];
    if (nd >= 3 && d.bz)
        for (int i = tid; i < Nx * Ny * F2; i += nt) ((T*)d.bz)[(size_t)n * Nx * Ny * F2 + i] = FZ[i];
}

// Energy outflow of the closure from the angular-flux moments of a pass (no sweep): one thread per
// (direction, cell, energy-face moment).
__global__ __launch_bounds__(256) void bfp_out_kernel(const PassDesc* __restrict__ pds, long n) {
    const PassDesc& d = pds[0];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int FE = d.fm_e, KE = d.ho_KE, NM = d.ho_K;
    const long Ncell = (long)d.Nx * d.Ny * d.Nz;
    const long mm = i % FE, r = i / FE, cell = r % Ncell, dir = r / Ncell;
    const double* psi = (const double*)d.psi + (size_t)dir * d.Nc + (size_t)NM * cell;
    double v = d.ho_ome[0] * (d.psiE_in ? d.psiE_in[i] : 0.0);
    for (int e = 0; e < KE; ++e) {
        const int j = d.emap[e * FE + mm];
        if (j >= 0) v += sqrt(2.0 * e + 1.0) * ((e & 1) ? -1.0 : 1.0) * d.ho_ome[e + 1] * psi[j];
    }
    d.psiE_out[i] = v;
}

template <typename T, int KE, int KX, int KY, int KZ, bool FC>
void bfp_launch(const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    auto fn = sweep_bfp_kernel<T, KE, KX, KY, KZ, FC>;
    static bool attr = false;
    if (!attr) {
        RT_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    if (a) hipExtLaunchKernelGGL(fn, dim3(Nd), dim3(256), (uint32_t)lds, st, a, b, 0u, d_pd);
    else hipLaunchKernelGGL(fn, dim3(Nd), dim3(256), (uint32_t)lds, st, d_pd);
}

template <typename T, int KE, int KX, int KY, int KZ>
void bfp_fc(bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    // partial coupling differs from full coupling only when two orders or more exceed 1; full
    // coupling is instantiated up to SWEEP_BFP_NM moments
    if constexpr ((KE > 1) + (KX > 1) + (KY > 1) + (KZ > 1) >= 2) {
        if (!fc) return bfp_launch<T, KE, KX, KY, KZ, false>(d_pd, Nd, lds, st);
    }
    if constexpr (KE * KX * KY * KZ <= SWEEP_BFP_NM) return bfp_launch<T, KE, KX, KY, KZ, true>(d_pd, Nd, lds, st);
    throw std::runtime_error("slowing-down cell with more than 8 moments");
}

template <typename T, int KE>
void bfp_dispatch_e(const PassDesc* d_pd, int Nd, int Kx, int Ky, int Kz, bool fc, size_t lds, hipStream_t st) {
    switch (Kx * 100 + Ky * 10 + Kz) {
        case 111: return bfp_fc<T, KE, 1, 1, 1>(fc, d_pd, Nd, lds, st);
        case 211: return bfp_fc<T, KE, 2, 1, 1>(fc, d_pd, Nd, lds, st);
        case 311: return bfp_fc<T, KE, 3, 1, 1>(fc, d_pd, Nd, lds, st);
        case 411: return bfp_fc<T, KE, 4, 1, 1>(fc, d_pd, Nd, lds, st);
        case 121: return bfp_fc<T, KE, 1, 2, 1>(fc, d_pd, Nd, lds, st);
        case 221: return bfp_fc<T, KE, 2, 2, 1>(fc, d_pd, Nd, lds, st);
        case 112: return bfp_fc<T, KE, 1, 1, 2>(fc, d_pd, Nd, lds, st);
        case 212: return bfp_fc<T, KE, 2, 1, 2>(fc, d_pd, Nd, lds, st);
        case 122: return bfp_fc<T, KE, 1, 2, 2>(fc, d_pd, Nd, lds, st);
        case 222: return bfp_fc<T, KE, 2, 2, 2>(fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("slowing-down closure orders out of range");
    }
}

}  // namespace

int bfp_cell_moments(int Ke, int Kx, int Ky, int Kz, bool fc) {
    return fc ? Ke * Kx * Ky * Kz : Ke + Kx + Ky + Kz - 3;
}

int bfp_face_moments(int axis, int Ke, int Kx, int Ky, int Kz, bool fc) {
    const int O[4] = {Ke, Kx, Ky, Kz};
    int prod = 1, sum = 0;
    for (int k = 0; k < 4; ++k)
        if (k != axis) { prod *= O[k]; sum += O[k] - 1; }
    return fc ? prod : 1 + sum;
}

bool sweep_bfp_supported(int ndims, int Ke, int Kx, int Ky, int Kz, bool fc) {
    if (Ke < 1 || Ke > 2 || Kx < 1 || Ky < 1 || Kz < 1) return false;
    if (ndims == 1 && (Kx > 4 || Ky > 1 || Kz > 1)) return false;
    if (ndims == 2 && (Kx > 2 || Ky > 2 || Kz > 1)) return false;
    if (ndims == 3 && (Kx > 2 || Ky > 2 || Kz > 2)) return false;
    return bfp_cell_moments(Ke, Kx, Ky, Kz, fc) <= SWEEP_BFP_NM;
}

size_t sweep_bfp_lds_bytes(int prec, int ndims, int Nx, int Ny, int Nz, int Ke, int Kx, int Ky, int Kz, bool fc) {
    const size_t f0 = bfp_face_moments(1, Ke, Kx, Ky, Kz, fc), f1 = bfp_face_moments(2, Ke, Kx, Ky, Kz, fc),
                 f2 = bfp_face_moments(3, Ke, Kx, Ky, Kz, fc);
    return (prec == P64 ? 8 : 4) * ((size_t)Ny * Nz * f0 + (ndims >= 2 ? (size_t)Nx * Nz * f1 : 0) +
                                    (ndims >= 3 ? (size_t)Nx * Ny * f2 : 0));
}

void launch_sweep_bfp(int prec, const PassDesc* d_pd, int Nd, int Ke, int Kx, int Ky, int Kz, bool fc, size_t lds,
                      hipStream_t st) {
    if (Nd <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("slowing-down closure face moments exceed the LDS");
    if (prec == P64) {
        if (Ke == 1) bfp_dispatch_e<double, 1>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
        else bfp_dispatch_e<double, 2>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
    } else {
        if (Ke == 1) bfp_dispatch_e<float, 1>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
        else bfp_dispatch_e<float, 2>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
    }
    RT_HIP(hipGetLastError());
}

void launch_bfp_out(const PassDesc* d_pd, long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(bfp_out_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_pd, n);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
