/*
 * Test infrastructure (oracle): a plain-C CPU restatement of the MiniFS
 * solver path, used only by tests/, __graft_entry__.smoke() and the CPU
 * leg of bench.py, and only as the checker or the CPU baseline.
 *
 * It restates, from the reference sources (paths relative to the MiniFS
 * tree): the FlowStar reader (src/Kernels/load_grid.cpp:129-314), SpecifyBC
 * (:857-902), ComputeMetrics (src/Base/grid_polyhedra.cpp:1071-1100 and
 * grid_polyhedra_auxiliary_functions.cpp:68-808), AllocAndInitFlowfield and
 * SetGhostVariables (load_grid.cpp:571-601, 953-1212), ComputeWeight3D_Node
 * (grid_polyhedra_auxiliary_functions.cpp:831-1049) and the driver loop of
 * src/main.cpp:198-291 with ZeroResiduals / InviscidFlux (cal_flux.cpp),
 * ComputeTimeStep (compute_timestep.cpp), ForwardLUSGS (forward_lusgs.cpp)
 * and CalculateGradient (cal_gradient.cpp).
 *
 * Mode 0 runs the reference's serial loops (face loops scattering into
 * cells, sequential sweeps).  Mode > 0 runs the same arithmetic with that
 * many OpenMP threads: per-cell gathers in ascending face order and
 * level-scheduled sweeps, which add the same terms in the same order.
 */
#ifndef MFS_ORACLE_H
#define MFS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_grid or_grid;

int or_load(const char* mesh_path, const char* bc_path, or_grid** out);
/* sets the flow condition (para_field_global.cpp:150-209) and runs the set-up */
int or_preprocess(or_grid* g, double mach, double alpha, double cfl);
/* runs iters iterations; hist receives iters*7 doubles
 * (5 residual norms, running dt_max, dt_min); t_loop the loop seconds */
int or_run(or_grid* g, int iters, int threads, double* hist, double* t_loop);
/* copies a named array: rho u v w p (nC+nB), res (5 nC), dt (nC), dq dqdx
 * dqdy dqdz (5 (nC+nB)), xfn..zfn area xfc..zfc (nF), xcc..zcc vol (nC+nB) */
int or_get(const or_grid* g, const char* name, double* dst, size_t count);
int or_sizes(const or_grid* g, int* nC, int* nF, int* nB, int* nN);
/* SetGhostVariables after each iteration's update (the call main.cpp:283
 * leaves commented out); off after or_preprocess */
int or_set_ghost_refresh(or_grid* g, int on);
/* overwrites a state array (rho u v w p, nC+nB) */
int or_set(or_grid* g, const char* name, const double* src, size_t count);
/* SetGhostVariables (load_grid.cpp:953-1212) on the current state */
int or_set_ghosts(or_grid* g);
/* the state limits of para_field_global.cpp:200-208 and the dt ratio limit
 * (:115): ratio_rhop_min (default 1e-4), ratio_rhop_max (10), ratio_p_break
 * (0.01), ratio_dtmax (1e80).  Set before or_preprocess for the set-up's
 * ghost clamps to see them; afterwards they apply from the next iteration. */
int or_set_limits(or_grid* g, double ratio_rhop_min, double ratio_rhop_max, double ratio_p_break, double ratio_dtmax);
/* Guard-branch counters: how many times (cells, or cell x stage) each
 * robustness branch was taken since or_load or the last reset. */
enum {
    OR_FWD_LIMIT1, OR_FWD_LIMIT2, OR_FWD_LIMIT3, /* DQ limiter stages, forward sweep (forward_lusgs.cpp:939-973) */
    OR_BWD_LIMIT1, OR_BWD_LIMIT2, OR_BWD_LIMIT3, /* backward sweep (:1184-1218) */
    OR_RELAX_RHO, OR_RELAX_P,                    /* CFL3D under-relaxation (:1406-1424) */
    OR_CLAMP_RHO_LO, OR_CLAMP_RHO_HI, OR_CLAMP_P_LO, OR_CLAMP_P_HI, /* update clamps (:1413-1427) */
    OR_CFL_RAMP, OR_CFL_PMIN, OR_CFL_DET,        /* LimitTimeStep (compute_timestep.cpp:886-902) */
    OR_DTMAX,                                    /* ratio_dtmax clamp (:929-945) */
    OR_GHOST_RHO_LO, OR_GHOST_RHO_HI, OR_GHOST_P_LO, OR_GHOST_P_HI, /* load_grid.cpp:1207-1210 */
    OR_NCOUNT
};
int or_counts(or_grid* g, long long* dst, int n, int reset);
void or_free(or_grid* g);
const char* or_error(void);

/* Point functions for property tests.  derived[] receives u_s v_s w_s
 * rho_s p_bar p_stag e_stag p_min p_max rho_min rho_max p_break e_stag_max
 * gamm1 for the given flow condition (para_field_global.cpp:150-209). */
void or_derived(double mach, double alpha, double derived[14]);
/* RoeFlux_noprec for one face (cal_flux.cpp:682-833) */
void or_roe_flux(const double ql[5], const double qr[5], double nx, double ny, double nz, double area, double mach,
                 double alpha, double f[5]);
/* FluxLUSGS3D (forward_lusgs.cpp:1271-1324) */
void or_flux_lusgs(const double q[5], const double dq[5], double nx, double ny, double nz, double mach, double alpha,
                   double f[5]);

#ifdef __cplusplus
}
#endif
#endif
