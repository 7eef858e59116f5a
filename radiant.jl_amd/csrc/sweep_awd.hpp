// sweep_awd.hpp -- sweeps with the adaptive weighted diamond closure (SN scheme "AWD",
// particle_transport/adaptive.jl): FP64, one direction per lane (1D) or per workgroup (2D/3D).
#pragma once
#include "kernels.hpp"

namespace rt {

// Orders with an adaptive closure: 1 and 2 in 1D (AWD0, AWD1), 1 per axis in 2D and 3D (AWD0);
// with slowing down, 1 on every axis and in energy.
bool sweep_awd_supported(int ndims, int Kx, int Ky, int Kz);
// One sweep of every direction of the pass in d_pd (FP64 psi, bin/bx faces as the general-closure
// walkers take them); K = spatial moments per cell (1D: the order; 2D/3D: 1); lds = the face bytes
// of a 2D/3D workgroup (sweep_2d_ho_lds_bytes / sweep_3d_ho_lds_bytes with orders 1; 0 when the
// 3D faces live in PassDesc::hoface).
void launch_sweep_awd(const PassDesc* d_pd, int Nd, int ndims, int K, size_t lds, hipStream_t st);
// The slowing-down types (csd) of order 1 on every axis, space and energy: the general slowing-down
// walker's faces and energy faces (lds = sweep_bfp_lds_bytes with orders 1, 0 with hoface);
// inv_dE = 1/ΔE of the group, the energy axis' h in the adaptive estimate.
void launch_sweep_bfp_awd(const PassDesc* d_pd, int Nd, double inv_dE, size_t lds, hipStream_t st);
// The 1D slowing-down cell of orders (2, 2) in (x, E) with full coupling and the coupled linear AWD1:
// one lane per direction, the x face [Nd][2] and the energy faces [Nd][cells][2].
void launch_sweep_1d_bfp_awd1(const PassDesc* d_pd, int Nd, double inv_dE, hipStream_t st);

}  // namespace rt
