// sweep_hog.hip -- 2D/3D sweeps of a general closure of any order up to HOG_KMAX per axis, for gfx950.
//
// The walkers of kernels.hip hold a cell's Nm x Nm system in registers, which bounds them at 12
// moments per cell.  Beyond that (sn_2D_bte.jl:35-116 and sn_3D_bte.jl:34-152 take any order) this
// walker solves the same cell system in a per-thread slice of device scratch (Nm^2 + Nm values),
// with partial pivoting by row swaps: one wave per direction walks the diagonal planes
// kx + ky (+ kz) = d in sweep order, one cell per lane, the face moments of every pencil in LDS (3D
// grids whose faces exceed it: in the direction's slice of hoface).  Cell moments (a, b, c): KX*KY*c +
// KX*b + a with full coupling; without it one index non-zero, at a + b + c (+ KX-1 when b > 0,
// + KX-1+KY-1 when c > 0); faces carry the moments over the other two axes by the same rule.  2D is
// the 3D walk with one z layer and no z streaming.  The cell system and the face updates are those of
// sweep_3d_ho_kernel.
#include "sweep_hog.hpp"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>

namespace rt {

namespace {
struct Hog {
    int KX, KY, KZ;
    bool fc;
    __device__ __host__ int nm() const { return fc ? KX * KY * KZ : KX + KY + KZ - 2; }
    __device__ bool valid(int a, int b, int c) const { return fc || (a > 0) + (b > 0) + (c > 0) <= 1; }
    __device__ int idx(int a, int b, int c) const {
        return fc ? KX * KY * c + KX * b + a : a + b + c + (b > 0 ? KX - 1 : 0) + (c > 0 ? KX - 1 + KY - 1 : 0);
    }
    __device__ int fx(int b, int c) const { return fc ? KY * c + b : b + c + (c > 0 ? KY - 1 : 0); }
    __device__ int fy(int a, int c) const { return fc ? KX * c + a : a + c + (c > 0 ? KX - 1 : 0); }
    __device__ int fz(int a, int b) const { return fc ? KX * b + a : a + b + (b > 0 ? KX - 1 : 0); }
    __device__ __host__ int f0() const { return fc ? KY * KZ : KY + KZ - 1; }
    __device__ __host__ int f1() const { return fc ? KX * KZ : KX + KZ - 1; }
    __device__ __host__ int f2() const { return fc ? KX * KY : KX + KY - 1; }
};

template <typename T>
__global__ __launch_bounds__(HOG_THREADS) void sweep_hog_kernel(const PassDesc* __restrict__ pds, int nd, Hog H,
                                                                 T* __restrict__ scratch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny, Nz = nd >= 3 ? d.Nz : 1, Nc = d.Nc;
    const int NM = H.nm(), F0 = H.f0(), F1 = H.f1(), F2 = H.f2();
    const size_t nface = (size_t)Ny * Nz * F0 + (size_t)Nx * Nz * F1 + (nd >= 3 ? (size_t)Nx * Ny * F2 : 0);
    T* FX = d.hoface ? (T*)d.hoface + (size_t)n * nface : (T*)smem;  // [Ny*Nz][F0]
    T* FY = FX + (size_t)Ny * Nz * F0;                               // [Nx*Nz][F1]
    T* FZ = FY + (size_t)Nx * Nz * F1;                               // [Nx*Ny][F2] (3D)
    const double om[3] = {d.mu[n], d.eta[n], nd >= 3 ? d.xi[n] : 0.0};
    const int K3[3] = {H.KX, H.KY, H.KZ};
    const double* ws[3] = {d.ho_om, d.ho_omy, d.ho_omz};
    T s[3], w[3][HOG_KMAX + 1], Cc[HOG_KMAX];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a] = om[a] > 0 ? T(1) : (om[a] < 0 ? T(-1) : T(0));
        for (int k = 0; k <= HOG_KMAX; ++k) w[a][k] = (k <= K3[a] && (a < 2 || nd >= 3)) ? (T)ws[a][k] : T(0);
    }
    for (int k = 0; k < HOG_KMAX; ++k) Cc[k] = (T)sqrt(2.0 * k + 1.0);
    auto spw = [](T v, int e) {
        T r = T(1);
        for (int k = 0; k < e; ++k) r *= v;
        return r;
    };
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* S = scratch + ((size_t)n * HOG_THREADS + tid) * ((size_t)NM * NM + NM);
    T* q = S + (size_t)NM * NM;
    for (int i = tid; i < Ny * Nz * F0; i += nt) FX[i] = d.bin_x ? (T)d.bin_x[(size_t)n * Ny * Nz * F0 + i] : T(0);
    for (int i = tid; i < Nx * Nz * F1; i += nt) FY[i] = d.bin_y ? (T)d.bin_y[(size_t)n * Nx * Nz * F1 + i] : T(0);
    if (nd >= 3)
        for (int i = tid; i < Nx * Ny * F2; i += nt) FZ[i] = d.bin_z ? (T)d.bin_z[(size_t)n * Nx * Ny * F2 + i] : T(0);
    __syncthreads();
    const bool fw[3] = {om[0] >= 0, om[1] >= 0, om[2] >= 0};
    for (int dg = 0; dg < Nx + Ny + Nz - 2; ++dg) {
        for (int i = tid; i < Nx * Ny; i += nt) {
            const int kx = i % Nx, ky = i / Nx, kz = dg - kx - ky;
            if (kz < 0 || kz >= Nz) continue;
            const int ix = fw[0] ? kx : Nx - 1 - kx, iy = fw[1] ? ky : Ny - 1 - ky, iz = fw[2] ? kz : Nz - 1 - kz;
            const size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
            const T h[3] = {(T)(fabs(om[0]) / d.dx[ix]), (T)(fabs(om[1]) / d.dy[iy]),
                            nd >= 3 ? (T)(fabs(om[2]) / d.dz[iz]) : T(0)};
            const T St = (T)d.sigt[d.mat[c * NM]];
            T* fx = FX + ((size_t)iy + (size_t)Ny * iz) * F0;
            T* fy = FY + ((size_t)ix + (size_t)Nx * iz) * F1;
            T* fz = nd >= 3 ? FZ + ((size_t)ix + (size_t)Nx * iy) * F2 : nullptr;
            for (int k = 0; k < NM * NM; ++k) S[k] = T(0);
            for (int k = 0; k < NM; ++k) S[(size_t)k * NM + k] = St;
            for (int a = 0; a < H.KX; ++a)
                for (int b = 0; b < H.KY; ++b)
                    for (int cz = 0; cz < H.KZ; ++cz) {
                        if (!H.valid(a, b, cz)) continue;
                        const int ii = H.idx(a, b, cz);
                        const int ia[3] = {a, b, cz};
                        for (int a2 = 0; a2 < H.KX; ++a2)
                            for (int b2 = 0; b2 < H.KY; ++b2)
                                for (int c2 = 0; c2 < H.KZ; ++c2) {
                                    if (!H.valid(a2, b2, c2)) continue;
                                    const int jj = H.idx(a2, b2, c2);
                                    const int jb[3] = {a2, b2, c2};
                                    T e = S[(size_t)ii * NM + jj];
                                    for (int ax = 0; ax < 3; ++ax) {
                                        const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                                        if (ia[o1] != jb[o1] || ia[o2] != jb[o2]) continue;
                                        const int u = ia[ax], v = jb[ax];
                                        if (u > v && ((u - v) & 1)) e -= T(2) * Cc[u] * h[ax] * s[ax] * Cc[v];
                                        e += Cc[u] * h[ax] * spw(s[ax], u) * Cc[v] * spw(s[ax], v) * w[ax][v + 1];
                                    }
                                    S[(size_t)ii * NM + jj] = e;
                                }
                        T r = psi[(size_t)ii + NM * c];
                        r -= Cc[a] * h[0] * (spw(s[0], a) * w[0][0] - spw(-s[0], a)) * fx[H.fx(b, cz)];
                        r -= Cc[b] * h[1] * (spw(s[1], b) * w[1][0] - spw(-s[1], b)) * fy[H.fy(a, cz)];
                        if (nd >= 3) r -= Cc[cz] * h[2] * (spw(s[2], cz) * w[2][0] - spw(-s[2], cz)) * fz[H.fz(a, b)];
                        q[ii] = r;
                    }
            // Gaussian elimination with partial pivoting
            for (int p = 0; p < NM; ++p) {
                int piv = p;
                for (int r2 = p + 1; r2 < NM; ++r2)
                    if (fabs(S[(size_t)r2 * NM + p]) > fabs(S[(size_t)piv * NM + p])) piv = r2;
                if (piv != p) {
                    for (int j = 0; j < NM; ++j) {
                        const T t = S[(size_t)p * NM + j];
                        S[(size_t)p * NM + j] = S[(size_t)piv * NM + j];
                        S[(size_t)piv * NM + j] = t;
                    }
                    const T t = q[p];
                    q[p] = q[piv];
                    q[piv] = t;
                }
                const T dp = S[(size_t)p * NM + p];
                for (int r2 = p + 1; r2 < NM; ++r2) {
                    const T f = S[(size_t)r2 * NM + p] / dp;
                    if (f == T(0)) continue;
                    for (int j = p + 1; j < NM; ++j) S[(size_t)r2 * NM + j] -= f * S[(size_t)p * NM + j];
                    q[r2] -= f * q[p];
                }
            }
            for (int r2 = NM - 1; r2 >= 0; --r2) {
                T v = q[r2];
                for (int j = r2 + 1; j < NM; ++j) v -= S[(size_t)r2 * NM + j] * q[j];
                q[r2] = v / S[(size_t)r2 * NM + r2];
            }
            for (int k = 0; k < NM; ++k) psi[(size_t)k + NM * c] = q[k];
            // outgoing faces, in place: each face moment starts from w_0 times itself
            for (int k = 0; k < F0; ++k) fx[k] *= w[0][0];
            for (int k = 0; k < F1; ++k) fy[k] *= w[1][0];
            if (nd >= 3)
                for (int k = 0; k < F2; ++k) fz[k] *= w[2][0];
            for (int a = 0; a < H.KX; ++a)
                for (int b = 0; b < H.KY; ++b)
                    for (int cz = 0; cz < H.KZ; ++cz) {
                        if (!H.valid(a, b, cz)) continue;
                        const T v = q[H.idx(a, b, cz)];
                        fx[H.fx(b, cz)] += Cc[a] * spw(s[0], a) * w[0][a + 1] * v;
                        fy[H.fy(a, cz)] += Cc[b] * spw(s[1], b) * w[1][b + 1] * v;
                        if (nd >= 3) fz[H.fz(a, b)] += Cc[cz] * spw(s[2], cz) * w[2][cz + 1] * v;
                    }
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * Nz * F0; i += nt) ((T*)d.bx)[(size_t)n * Ny * Nz * F0 + i] = FX[i];
    if (d.by)
        for (int i = tid; i < Nx * Nz * F1; i += nt) ((T*)d.by)[(size_t)n * Nx * Nz * F1 + i] = FY[i];
    if (nd >= 3 && d.bz)
        for (int i = tid; i < Nx * Ny * F2; i += nt) ((T*)d.bz)[(size_t)n * Nx * Ny * F2 + i] = FZ[i];
}

template <typename T>
void hog_launch(const PassDesc* d_pd, int Nd, int nd, Hog H, size_t lds, T* scratch, hipStream_t st) {
    auto fn = sweep_hog_kernel<T>;
    static bool attr = false;
    if (!attr) {
        const hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
        attr = true;
    }
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    if (a) hipExtLaunchKernelGGL(fn, dim3(Nd), dim3(HOG_THREADS), (uint32_t)lds, st, a, b, 0u, d_pd, nd, H, scratch);
    else hipLaunchKernelGGL(fn, dim3(Nd), dim3(HOG_THREADS), (uint32_t)lds, st, d_pd, nd, H, scratch);
}

// The slowing-down cell over (E, x, y, z), orders O = (KE, KX, KY, KZ): with full coupling moment
// (e, a, b, c) sits at e + KE*(a + KX*(b + KY*c)), energy fastest; without it only one index is
// non-zero and sits at that index plus the (order - 1) of the axes before it; a face carries the
// moments over the other three axes by the same rule.  The cell system, with C_k = sqrt(2k+1),
// h_a = |Ω_a|/Δ_a, s_a = sign(Ω_a), the stopping powers S⁻, S⁺ at the group's bounds and
// Sv = ((S⁻+S⁺)/2, (S⁻-S⁺)/(2√3)) (sn_1D_bfp.jl:44-125, sn_2D_bfp.jl:43-154, sn_3D_bfp.jl:51-191):
//   S_ij = St δ_ij + Σ_axes [other indices equal] (-[u>v, u-v odd] 2 C_u h s C_v + C_u h s^u C_v s^v w_(v+1))
//        + [spatial indices equal] ([e_i = 1] 2 C_1 Sv_(e_j) + C_ei S⁺ (-1)^ei C_ej (-1)^ej wE_(ej+1))
//   Q_j  = Qn_j - Σ_axes C_v h (s^v w_0 - (-s)^v) f_axis - C_e ((-1)^e S⁺ wE_0 - S⁻) f_E
// and the outgoing faces f <- w_0 f + Σ_v C_v s^v w_(v+1) ψ (energy: (-1)^e for s^v).  The in-group
// energy term takes that closed form for energy orders up to 2.  In 1D a direction with
// |μ| <= 1e-10 does not stream and keeps its x face (sn_1D_bfp.jl:52-54).
struct Hog4 {
    int O[4];
    bool fc;
    __device__ __host__ int nm() const { return fc ? O[0] * O[1] * O[2] * O[3] : O[0] + O[1] + O[2] + O[3] - 3; }
    __device__ __host__ int face(int skip) const {
        int prod = 1, sum = 0;
        for (int k = 0; k < 4; ++k)
            if (k != skip) { prod *= O[k]; sum += O[k] - 1; }
        return fc ? prod : 1 + sum;
    }
    __device__ bool valid(const int* m) const { return fc || (m[0] > 0) + (m[1] > 0) + (m[2] > 0) + (m[3] > 0) <= 1; }
    __device__ int index(const int* m, int skip) const {
        int r = 0, stride = 1, off = 0;
        for (int k = 0; k < 4; ++k) {
            if (k == skip) continue;
            if (fc) {
                r += m[k] * stride;
                stride *= O[k];
            } else {
                if (m[k] > 0) r = m[k] + off;
                off += O[k] - 1;
            }
        }
        return r;
    }
    // the multi-index of moment slot t (0 <= t < O0*O1*O2*O3, energy fastest)
    __device__ void unpack(int t, int* m) const {
        for (int k = 0; k < 4; ++k) {
            m[k] = t % O[k];
            t /= O[k];
        }
    }
    __device__ int slots() const { return O[0] * O[1] * O[2] * O[3]; }
};

template <typename T>
__global__ __launch_bounds__(HOG_THREADS) void sweep_hog_bfp_kernel(const PassDesc* __restrict__ pds, Hog4 H,
                                                                     T* __restrict__ scratch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int nd = d.ndims, Nx = d.Nx, Ny = nd >= 2 ? d.Ny : 1, Nz = nd >= 3 ? d.Nz : 1, Nc = d.Nc;
    const size_t Ncell = (size_t)Nx * Ny * Nz;
    const int NM = H.nm(), FE = H.face(0), F0 = H.face(1), F1 = H.face(2), F2 = H.face(3), NS = H.slots();
    const size_t nface = (size_t)Ny * Nz * F0 + (nd >= 2 ? (size_t)Nx * Nz * F1 : 0) + (nd >= 3 ? (size_t)Nx * Ny * F2 : 0);
    T* FX = d.hoface ? (T*)d.hoface + (size_t)n * nface : (T*)smem;  // [Ny*Nz][F0]
    T* FY = FX + (size_t)Ny * Nz * F0;                                // [Nx*Nz][F1] (2D, 3D)
    T* FZ = FY + (nd >= 2 ? (size_t)Nx * Nz * F1 : 0);               // [Nx*Ny][F2] (3D)
    const double om[3] = {d.mu[n], nd >= 2 ? d.eta[n] : 0.0, nd >= 3 ? d.xi[n] : 0.0};
    const bool tang = nd == 1 && fabs(om[0]) <= 1e-10;
    const double* ws[3] = {d.ho_om, d.ho_omy, d.ho_omz};
    T s[3], w[3][HOG_KMAX + 1], we[3], Cc[HOG_KMAX];
    for (int a = 0; a < 3; ++a) {
        s[a] = om[a] > 0 ? T(1) : (om[a] < 0 ? T(-1) : T(0));
        if (a == 0 && tang) s[a] = T(1);
        for (int k = 0; k <= HOG_KMAX; ++k) w[a][k] = (k <= H.O[a + 1] && a < nd) ? (T)ws[a][k] : T(0);
    }
    for (int k = 0; k < 3; ++k) we[k] = k <= H.O[0] ? (T)d.ho_ome[k] : T(0);
    for (int k = 0; k < HOG_KMAX; ++k) Cc[k] = (T)sqrt(2.0 * k + 1.0);
    auto spw = [](T v, int e) {
        T r = T(1);
        for (int k = 0; k < e; ++k) r *= v;
        return r;
    };
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    const double* Ein = d.psiE_in ? d.psiE_in + (size_t)n * Ncell * FE : nullptr;
    double* Eout = d.psiE_out ? d.psiE_out + (size_t)n * Ncell * FE : nullptr;
    // per lane: the cell system S [NM][NM] and its right-hand side; then the energy face [FE]
    T* S = scratch + ((size_t)n * HOG_THREADS + tid) * ((size_t)NM * NM + NM);
    T* q = S + (size_t)NM * NM;
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
            const int mt = d.mat[c * NM];
            const T St = (T)d.sigt[mt];
            const T Sm = (T)d.csdS[2 * mt], Sp = (T)d.csdS[2 * mt + 1];
            const T Sv[2] = {(Sm + Sp) / T(2), (Sm - Sp) / (T(2) * sq3)};
            T* f[3] = {FX + ((size_t)iy + (size_t)Ny * iz) * F0, nd >= 2 ? FY + ((size_t)ix + (size_t)Nx * iz) * F1 : nullptr,
                       nd >= 3 ? FZ + ((size_t)ix + (size_t)Nx * iy) * F2 : nullptr};
            for (int k = 0; k < NM * NM; ++k) S[k] = T(0);
            for (int k = 0; k < NM; ++k) S[(size_t)k * NM + k] = St;
            for (int t = 0; t < NS; ++t) {
                int mi[4];
                H.unpack(t, mi);
                if (!H.valid(mi)) continue;
                const int ii = H.index(mi, -1);
                for (int t2 = 0; t2 < NS; ++t2) {
                    int mj[4];
                    H.unpack(t2, mj);
                    if (!H.valid(mj)) continue;
                    const int jj = H.index(mj, -1);
                    T e = S[(size_t)ii * NM + jj];
                    for (int ax = 0; ax < 3; ++ax) {
                        const int o1 = 1 + (ax + 1) % 3, o2 = 1 + (ax + 2) % 3;
                        if (mi[0] != mj[0] || mi[o1] != mj[o1] || mi[o2] != mj[o2]) continue;
                        const int u = mi[ax + 1], v = mj[ax + 1];
                        if (u > v && ((u - v) & 1)) e -= T(2) * Cc[u] * h[ax] * s[ax] * Cc[v];
                        e += Cc[u] * h[ax] * spw(s[ax], u) * Cc[v] * spw(s[ax], v) * w[ax][v + 1];
                    }
                    if (mi[1] == mj[1] && mi[2] == mj[2] && mi[3] == mj[3]) {
                        if (mi[0] == 1) e += T(2) * Cc[1] * Sv[mj[0]];
                        e += Cc[mi[0]] * Sp * spw(T(-1), mi[0]) * Cc[mj[0]] * spw(T(-1), mj[0]) * we[mj[0] + 1];
                    }
                    S[(size_t)ii * NM + jj] = e;
                }
                T r = psi[(size_t)ii + NM * c];
                for (int ax = 0; ax < nd; ++ax) {
                    const int u = mi[ax + 1];
                    r -= Cc[u] * h[ax] * (spw(s[ax], u) * w[ax][0] - spw(-s[ax], u)) * f[ax][H.index(mi, ax + 1)];
                }
                const T fe = Ein ? (T)Ein[c * FE + H.index(mi, 0)] : T(0);
                r -= Cc[mi[0]] * (spw(T(-1), mi[0]) * Sp * we[0] - Sm) * fe;
                q[ii] = r;
            }
            // Gaussian elimination with partial pivoting
            for (int p = 0; p < NM; ++p) {
                int piv = p;
                for (int r2 = p + 1; r2 < NM; ++r2)
                    if (fabs(S[(size_t)r2 * NM + p]) > fabs(S[(size_t)piv * NM + p])) piv = r2;
                if (piv != p) {
                    for (int j = 0; j < NM; ++j) {
                        const T t = S[(size_t)p * NM + j];
                        S[(size_t)p * NM + j] = S[(size_t)piv * NM + j];
                        S[(size_t)piv * NM + j] = t;
                    }
                    const T t = q[p];
                    q[p] = q[piv];
                    q[piv] = t;
                }
                const T dp = S[(size_t)p * NM + p];
                for (int r2 = p + 1; r2 < NM; ++r2) {
                    const T fct = S[(size_t)r2 * NM + p] / dp;
                    if (fct == T(0)) continue;
                    for (int j = p + 1; j < NM; ++j) S[(size_t)r2 * NM + j] -= fct * S[(size_t)p * NM + j];
                    q[r2] -= fct * q[p];
                }
            }
            for (int r2 = NM - 1; r2 >= 0; --r2) {
                T v = q[r2];
                for (int j = r2 + 1; j < NM; ++j) v -= S[(size_t)r2 * NM + j] * q[j];
                q[r2] = v / S[(size_t)r2 * NM + r2];
            }
            for (int k = 0; k < NM; ++k) psi[(size_t)k + NM * c] = q[k];
            // energy outflow (from the inflow still in Ein), then the spatial faces in place
            if (Eout) {
                T* oe = S;  // the eliminated system is no longer needed: FE <= NM * NM slots
                for (int k = 0; k < FE; ++k) oe[k] = we[0] * (Ein ? (T)Ein[c * FE + k] : T(0));
                for (int t = 0; t < NS; ++t) {
                    int mi[4];
                    H.unpack(t, mi);
                    if (!H.valid(mi)) continue;
                    oe[H.index(mi, 0)] += Cc[mi[0]] * spw(T(-1), mi[0]) * we[mi[0] + 1] * q[H.index(mi, -1)];
                }
                for (int k = 0; k < FE; ++k) Eout[c * FE + k] = (double)oe[k];
            }
            const int Fa[3] = {F0, F1, F2};
            for (int ax = 0; ax < nd; ++ax) {
                if (ax == 0 && tang) continue;
                for (int k = 0; k < Fa[ax]; ++k) f[ax][k] *= w[ax][0];
                for (int t = 0; t < NS; ++t) {
                    int mi[4];
                    H.unpack(t, mi);
                    if (!H.valid(mi)) continue;
                    const int u = mi[ax + 1];
                    f[ax][H.index(mi, ax + 1)] += Cc[u] * spw(s[ax], u) * w[ax][u + 1] * q[H.index(mi, -1)];
                }
            }
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

template <typename T>
void hog_bfp_launch(const PassDesc* d_pd, int Nd, Hog4 H, size_t lds, T* scratch, hipStream_t st) {
    auto fn = sweep_hog_bfp_kernel<T>;
    static bool attr = false;
    if (!attr) {
        const hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
        attr = true;
    }
    hipEvent_t a, b;
    take_launch_events(&a, &b);
    if (a) hipExtLaunchKernelGGL(fn, dim3(Nd), dim3(HOG_THREADS), (uint32_t)lds, st, a, b, 0u, d_pd, H, scratch);
    else hipLaunchKernelGGL(fn, dim3(Nd), dim3(HOG_THREADS), (uint32_t)lds, st, d_pd, H, scratch);
}
}  // namespace

bool sweep_hog_supported(int ndims, int Kx, int Ky, int Kz, bool fc) {
    if (ndims < 2 || Kx < 1 || Ky < 1 || Kz < 1 || Kx > HOG_KMAX || Ky > HOG_KMAX || Kz > HOG_KMAX) return false;
    if (ndims == 2 && Kz != 1) return false;
    const int nm = fc ? Kx * Ky * Kz : Kx + Ky + Kz - 2;
    return nm <= HOG_NM_MAX;
}

size_t sweep_hog_scratch(int Nd, int Kx, int Ky, int Kz, bool fc) {
    const size_t nm = fc ? (size_t)Kx * Ky * Kz : (size_t)(Kx + Ky + Kz - 2);
    return (size_t)Nd * HOG_THREADS * (nm * nm + nm);
}

void launch_sweep_hog(int prec, const PassDesc* d_pd, int Nd, int ndims, int Kx, int Ky, int Kz, bool fc, size_t lds,
                      double* scratch, hipStream_t st) {
    if (Nd <= 0) return;
    if (!sweep_hog_supported(ndims, Kx, Ky, Kz, fc)) throw std::runtime_error("general closure orders out of range");
    if (lds > 160 * 1024) throw std::runtime_error("general closure face moments exceed the LDS");
    const Hog H{Kx, Ky, Kz, fc};
    if (prec == P64) hog_launch<double>(d_pd, Nd, ndims, H, lds, scratch, st);
    else hog_launch<float>(d_pd, Nd, ndims, H, lds, (float*)scratch, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}

bool sweep_hog_bfp_supported(int ndims, int Ke, int Kx, int Ky, int Kz, bool fc) {
    if (ndims < 1 || ndims > 3 || Ke < 1 || Ke > 2 || Kx < 1 || Ky < 1 || Kz < 1) return false;
    if (Kx > HOG_KMAX || Ky > HOG_KMAX || Kz > HOG_KMAX || (ndims < 2 && Ky != 1) || (ndims < 3 && Kz != 1)) return false;
    return (fc ? Ke * Kx * Ky * Kz : Ke + Kx + Ky + Kz - 3) <= HOG_NM_MAX;
}

size_t sweep_hog_bfp_scratch(int Nd, int Ke, int Kx, int Ky, int Kz, bool fc) {
    const size_t nm = fc ? (size_t)Ke * Kx * Ky * Kz : (size_t)(Ke + Kx + Ky + Kz - 3);
    return (size_t)Nd * HOG_THREADS * (nm * nm + nm);
}

void launch_sweep_hog_bfp(int prec, const PassDesc* d_pd, int Nd, int ndims, int Ke, int Kx, int Ky, int Kz, bool fc,
                          size_t lds, double* scratch, hipStream_t st) {
    if (Nd <= 0) return;
    if (!sweep_hog_bfp_supported(ndims, Ke, Kx, Ky, Kz, fc))
        throw std::runtime_error("general slowing-down closure orders out of range");
    if (lds > 160 * 1024) throw std::runtime_error("general slowing-down closure face moments exceed the LDS");
    const Hog4 H{{Ke, Kx, Ky, Kz}, fc};
    if (prec == P64) hog_bfp_launch<double>(d_pd, Nd, H, lds, scratch, st);
    else hog_bfp_launch<float>(d_pd, Nd, H, lds, (float*)scratch, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}

}  // namespace rt
