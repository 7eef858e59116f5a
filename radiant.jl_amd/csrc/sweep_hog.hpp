// sweep_hog.hpp -- 2D/3D sweeps of a general closure whose cell system does not fit the register
// walkers of kernels.hip (more than 12 moments per cell, or orders above 4 in 2D / 3 in 3D).
#pragma once
#include "kernels.hpp"

namespace rt {

constexpr int HOG_KMAX = 8;     // order per axis at most
constexpr int HOG_NM_MAX = 64;  // moments per cell at most
constexpr int HOG_THREADS = 64; // one wave per direction

// Orders the general walker takes (1..HOG_KMAX per axis, at most HOG_NM_MAX moments per cell).
bool sweep_hog_supported(int ndims, int Kx, int Ky, int Kz, bool fc);
// Doubles of cell-system scratch the walker needs for Nd directions (a per-thread slice of
// NM^2 + NM values).
size_t sweep_hog_scratch(int Nd, int Kx, int Ky, int Kz, bool fc);
// One sweep of every direction of the pass (psi, faces and bin/bx as the 2D/3D general-closure
// walkers take them); lds = their face bytes (0 with PassDesc::hoface in 3D).
void launch_sweep_hog(int prec, const PassDesc* d_pd, int Nd, int ndims, int Kx, int Ky, int Kz, bool fc, size_t lds,
                      double* scratch, hipStream_t st);

// The slowing-down cell (E, x, y, z) of the same walk (energy order 1-2, spatial orders up to HOG_KMAX
// per axis, at most HOG_NM_MAX moments): the cells beyond sweep_bfp.hip's register walker.
bool sweep_hog_bfp_supported(int ndims, int Ke, int Kx, int Ky, int Kz, bool fc);
// its scratch, in doubles
size_t sweep_hog_bfp_scratch(int Nd, int Ke, int Kx, int Ky, int Kz, bool fc);
void launch_sweep_hog_bfp(int prec, const PassDesc* d_pd, int Nd, int ndims, int Ke, int Kx, int Ky, int Kz, bool fc,
                          size_t lds, double* scratch, hipStream_t st);

}  // namespace rt
