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

}  // namespace rt
