/*
 * mfs_gpu_shim.h -- the two calls a MiniFS driver adds when it links
 * mfs_gpu_shim.cpp (the reference's solver entry points on libmfs_hip.so)
 * in place of src/Kernels/cal_flux.cpp, compute_timestep.cpp,
 * forward_lusgs.cpp and cal_gradient.cpp.
 */
#ifndef MFS_GPU_SHIM_H
#define MFS_GPU_SHIM_H

#include "grid_polyhedra.h"

/* Copies the device results into the PolyGrid's host arrays: rho, u, v, w,
 * p (nTCell+nBFace each), res (5 nTCell), dt (nTCell), DQ and the gradients
 * (5 (nTCell+nBFace) each), allocating any that is still NULL, and sets the
 * globals dt_max, dt_min and cfl_min.  src/main.cpp calls it once after its
 * loop (:291), before the residual printout (:294-329). */
void MfsDownloadResults(PolyGrid* grid);

/* Sends the host state rho, u, v, w, p (ghosts included) to the device, for
 * a driver that changes the state after PreprocessGrid. */
void MfsUploadState(PolyGrid* grid);

#endif
