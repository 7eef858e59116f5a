// em_source.hip -- the source term of an external magnetic field (sn_inner_pass.jl:26-32), for gfx950.
//
// Before the moment -> direction contraction of a pass with a non-zero state, the active source
// rows become q_out[k][c] = q_in[k][c] + sum_j ℳ_EM[k][j] Φ[j][c] over the active moments (the
// moments ℳ_EM reads or writes are active), one thread per (moment, column), columns fastest: the
// state rows are read coalesced, ℳ_EM (Nact x Nact, row-major) from cache.  q_in null stands for
// zero (a homogeneous pass).
#include "em_source.hpp"

#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>

namespace rt {

__global__ __launch_bounds__(256) void em_source_kernel(const double* __restrict__ phi_in, const int* __restrict__ rows,
                                                        const double* __restrict__ Mem, const double* __restrict__ qin,
                                                        double* __restrict__ qout, int Nact, int Nc, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / Nc), c = (int)(i - (long)k * Nc);
        const double* __restrict__ Mk = Mem + (size_t)k * Nact;
        double s = 0.0;
        for (int j = 0; j < Nact; ++j) s += Mk[j] * phi_in[(size_t)(rows ? rows[j] : j) * Nc + c];
        qout[i] = (qin ? qin[i] : 0.0) + s;
    }
}

void launch_em_source(const double* phi_in, const int* rows, const double* Mem, const double* qin, double* qout,
                      int Nact, int Nc, hipStream_t st) {
    const long n = (long)Nact * Nc;
    if (n <= 0) return;
    const long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(em_source_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, st, phi_in,
                       rows, Mem, qin, qout, Nact, Nc, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}

}  // namespace rt
