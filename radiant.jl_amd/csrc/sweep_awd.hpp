// sweep_awd.hpp -- sweeps with the adaptive weighted diamond closure (SN scheme "AWD",
// particle_transport/adaptive.jl): FP64, one direction per lane (1D) or per workgroup (2D/3D).
#pragma once
#include "kernels.hpp"

namespace rt {

// Orders with an adaptive closure: 1 and 2 in 1D (AWD0, AWD1), 1 per axis in 2D and 3D (AWD0).
bool sweep_awd_supported(int ndims, int Kx, int Ky, int Kz);
// One sweep of every direction of the pass in d_pd (FP64 psi, bin/bx faces as the general-closure
// walkers take them); K = spatial moments per cell (1D: the order; 2D/3D: 1); lds = the face bytes
// of a 2D/3D workgroup (sweep_2d_ho_lds_bytes / sweep_3d_ho_lds_bytes with orders 1; 0 when the
// 3D faces live in PassDesc::hoface).
void launch_sweep_awd(const PassDesc* d_pd, int Nd, int ndims, int K, size_t lds, hipStream_t st);

}  // namespace rt
