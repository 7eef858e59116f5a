// em_source.hpp -- the source term of an external magnetic field, ℳ_EM Φ (sn_inner_pass.jl:26-32).
#pragma once
#include <hip/hip_runtime.h>

namespace rt {

// q_out[k][c] = q_in[k][c] + sum_j Mem[k][j] Φ[j][c] for the Nact active moments k, j and the Nc moment
// columns; Φ row j at phi_in + rows[j] * Nc (rows null: row j), q_in null: zero.
void launch_em_source(const double* phi_in, const int* rows, const double* Mem, const double* qin, double* qout,
                      int Nact, int Nc, hipStream_t st);

}  // namespace rt
