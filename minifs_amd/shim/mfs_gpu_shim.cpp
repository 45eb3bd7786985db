// mfs_gpu_shim.cpp -- the reference-side binding of libmfs_hip.so.
//
// A MiniFS tree compiles this file in place of src/Kernels/cal_flux.cpp,
// compute_timestep.cpp, forward_lusgs.cpp and cal_gradient.cpp and links
// with -lmfs_hip.  It defines the functions of those files that the rest of
// MiniFS calls, with the reference's names and signatures:
//
//   ZeroResiduals(PolyGrid*)                 cal_flux.h:24
//   InviscidFlux(PolyGrid*, RealFlow**, int) cal_flux.h:27
//   InviscidFlux_TDTree(...)                 cal_flux.h:41 (TDTREE builds)
//   ComputeTimeStep(PolyGrid*)               compute_timestep.h:24
//   ForwardLUSGS(PolyGrid*, int)             forward_lusgs.h:23
//   CalculateGradient(PolyGrid*)             cal_gradient.h:23
//   AllocAndCalQuantityGradient(PolyGrid*)   cal_gradient.h:22 (called by
//                                            PreprocessGrid, load_grid.cpp:531)
//
// The reference's LoadGrid, Decoupling, PreprocessGrid, PassInterfaceData
// and SetGhostVariables (src/Kernels/load_grid.cpp, src/Base/*) stay as
// they are; the device context is built from the arrays PreprocessGrid
// produced, the first time AllocAndCalQuantityGradient runs.  From then on
// the state lives on the device: the loop of src/main.cpp:198-291 runs
// unchanged, and MfsDownloadResults brings the results back to the
// PolyGrid when the driver reads them.  Errors are the reference's: a
// message on stderr and exit(1).
#include <cstdlib>
#include <iostream>
#include <vector>

#include "mfs.h"
#include "mfs_gpu_shim.h"

#include "cal_flux.h"
#include "cal_gradient.h"
#include "compute_timestep.h"
#include "forward_lusgs.h"
#include "memory_util.h"
#include "para_field_global.h"

// Defined in para_field_global.cpp (:200-202) without a declaration in its
// header.
extern RealFlow ratio_rhop_max, ratio_rhop_min, ratio_p_break;

namespace {

mfs_ctx* g_ctx = nullptr;  // one solve per process, as in main.cpp
std::vector<int> g_bctype;

void check(int rc, const char* what) {
    if (rc == MFS_OK) return;
    std::cerr << "libmfs_hip: " << what << ": " << mfs_last_error() << std::endl;
    exit(1);
}

// The solver globals of para_field_global.cpp as mfs_params.  The derived
// values are taken from the globals themselves, so the device runs with
// exactly the numbers the reference computed.
mfs_params params_from_globals() {
    mfs_params p;
    mfs_params_default(&p);
    p.mach = mach;
    p.alpha = alpha;
    p.beita = beita;
    p.re = re;
    p.T = T;
    p.gam = gam;
    p.gascon = gascon;
    p.cfl_start = cfl_start;
    p.cfl_end = cfl_end;
    p.cfl_nstep = cfl_nstep;
    p.epsa_r = epsa_r;
    p.lhs_omga = lhs_omga;
    p.ratio_dtmax = ratio_dtmax;
    p.ratio_rhop_max = ratio_rhop_max;
    p.ratio_rhop_min = ratio_rhop_min;
    p.ratio_p_break = ratio_p_break;
    check(mfs_params_derive(&p), "mfs_params_derive");
    p.amu = amu;
    p.ainf = ainf;
    p.uqq = uqq;
    p.u_s = u_s;
    p.v_s = v_s;
    p.w_s = w_s;
    p.p_s = p_s;
    p.rho_s = rho_s;
    p.p_bar = p_bar;
    p.gamm1 = gamm1;
    p.p_stag = p_stag;
    p.e_stag = e_stag;
    p.p_min = p_min;
    p.p_max = p_max;
    p.rho_min = rho_min;
    p.rho_max = rho_max;
    p.p_break = p_break;
    p.e_stag_max = e_stag_max;
    p.alf_l = alf_l;
    p.alf_n = alf_n;
    return p;
}

mfs_ctx* ctx_for(PolyGrid* grid) {
    if (g_ctx) return g_ctx;
    const mfs_params p = params_from_globals();
    const IntType nB = grid->GetNBFace();
    g_bctype.resize(nB);
    for (IntType k = 0; k < nB; ++k) g_bctype[k] = grid->Getbcr()[k]->GetType();
    mfs_mesh_view v;
    v.nTCell = grid->GetNTCell();
    v.nTFace = grid->GetNTFace();
    v.nBFace = nB;
    v.nTNode = grid->GetNTNode();
    v.x = grid->GetX();
    v.y = grid->GetY();
    v.z = grid->GetZ();
    v.nNPF = grid->GetnNPF();
    v.f2n = grid->Getf2n();
    v.f2c = grid->Getf2c();
    v.bc_type = g_bctype.data();
    v.xfn = grid->GetXfn();
    v.yfn = grid->GetYfn();
    v.zfn = grid->GetZfn();
    v.area = grid->GetFaceArea();
    v.xfc = grid->GetXfc();
    v.yfc = grid->GetYfc();
    v.zfc = grid->GetZfc();
    v.xcc = grid->GetXcc();
    v.ycc = grid->GetYcc();
    v.zcc = grid->GetZcc();
    v.vol = grid->GetCellVol();
    v.rho = grid->GetRho();
    v.u = grid->GetU();
    v.v = grid->GetV();
    v.w = grid->GetW();
    v.p = grid->GetP();
    check(mfs_ctx_create(&v, &p, MFS_FP64, &g_ctx), "mfs_ctx_create");
    return g_ctx;
}

RealFlow* host_array(RealFlow* cur, size_t n) {
    if (cur) return cur;
    RealFlow* a = nullptr;
    snew_array_1D(a, n);
    for (size_t i = 0; i < n; ++i) a[i] = 0.0;
    return a;
}

void download(int field, RealFlow* dst, size_t n, const char* what) { check(mfs_download(g_ctx, field, dst, n), what); }

}  // namespace

void ZeroResiduals(PolyGrid* grid) { check(mfs_zero_residuals(ctx_for(grid)), "ZeroResiduals"); }

void InviscidFlux(PolyGrid* grid, RealFlow** /*limit*/, IntType /*level*/) {
    check(mfs_inviscid_flux(ctx_for(grid)), "InviscidFlux");
}

#ifdef TDTREE
// The TDTree build's face loop gives the same per-cell sums on the
// TDTree-renumbered mesh (each cell's faces in ascending face id).
void InviscidFlux_TDTree(PolyGrid* grid, RealFlow** /*limit*/, IntType /*level*/) {
    check(mfs_inviscid_flux(ctx_for(grid)), "InviscidFlux_TDTree");
}
#endif

void ComputeTimeStep(PolyGrid* grid) { check(mfs_compute_timestep(ctx_for(grid)), "ComputeTimeStep"); }

void ForwardLUSGS(PolyGrid* grid, IntType /*level*/) { check(mfs_forward_lusgs(ctx_for(grid)), "ForwardLUSGS"); }

// main.cpp follows every CalculateGradient with iter_done++ (:287-289);
// the context's own count (it drives the CFL ramp, compute_timestep.cpp:
// 868-871) and its residual history advance here.
void CalculateGradient(PolyGrid* grid) {
    mfs_ctx* c = ctx_for(grid);
    check(mfs_calculate_gradient(c), "CalculateGradient");
    check(mfs_iteration_done(c), "CalculateGradient");
}

// cal_gradient.cpp:50-78: zeroed host gradients, then the gradient of the
// initial state (not an iteration).
void AllocAndCalQuantityGradient(PolyGrid* grid) {
    const size_t n = 5 * ((size_t)grid->GetNTCell() + grid->GetNBFace());
    grid->SetDqdx(host_array(nullptr, n));
    grid->SetDqdy(host_array(nullptr, n));
    grid->SetDqdz(host_array(nullptr, n));
    check(mfs_calculate_gradient(ctx_for(grid)), "AllocAndCalQuantityGradient");
}

void MfsDownloadResults(PolyGrid* grid) {
    mfs_ctx* c = ctx_for(grid);
    check(mfs_sync(c), "mfs_sync");
    const size_t nC = grid->GetNTCell(), n = nC + grid->GetNBFace();
    if (!grid->GetRes()) grid->SetRes(host_array(nullptr, 5 * nC));
    if (!grid->GetDt()) grid->SetDt(host_array(nullptr, nC));
    if (!grid->GetDQ()) grid->SetDQ(host_array(nullptr, 5 * n));
    download(MFS_F_RHO, grid->GetRho(), n, "download rho");
    download(MFS_F_U, grid->GetU(), n, "download u");
    download(MFS_F_V, grid->GetV(), n, "download v");
    download(MFS_F_W, grid->GetW(), n, "download w");
    download(MFS_F_P, grid->GetP(), n, "download p");
    download(MFS_F_RES, grid->GetRes(), 5 * nC, "download res");
    download(MFS_F_DT, grid->GetDt(), nC, "download dt");
    download(MFS_F_DQ, grid->GetDQ(), 5 * n, "download DQ");
    download(MFS_F_DQDX, grid->GetDqdx(), 5 * n, "download dqdx");
    download(MFS_F_DQDY, grid->GetDqdy(), 5 * n, "download dqdy");
    download(MFS_F_DQDZ, grid->GetDqdz(), 5 * n, "download dqdz");
    mfs_scalars s;
    check(mfs_get_scalars(c, &s), "mfs_get_scalars");
    dt_max = s.dt_max;
    dt_min = s.dt_min;
    cfl_min = s.cfl_min;
}

void MfsUploadState(PolyGrid* grid) {
    mfs_ctx* c = ctx_for(grid);
    const size_t n = (size_t)grid->GetNTCell() + grid->GetNBFace();
    check(mfs_upload(c, MFS_F_RHO, grid->GetRho(), n), "upload rho");
    check(mfs_upload(c, MFS_F_U, grid->GetU(), n), "upload u");
    check(mfs_upload(c, MFS_F_V, grid->GetV(), n), "upload v");
    check(mfs_upload(c, MFS_F_W, grid->GetW(), n), "upload w");
    check(mfs_upload(c, MFS_F_P, grid->GetP(), n), "upload p");
}
