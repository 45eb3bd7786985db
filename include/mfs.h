/*
 * mfs.h -- C ABI of libmfs_hip.so, the MI355X (gfx950) implementation of the
 * MiniFS per-iteration solver path: Roe inviscid flux with residual
 * assembly, local time step, LU-SGS forward/backward sweeps with the
 * flow-field update, and node-based Green-Gauss gradients.
 *
 * Reference interfaces replaced (paths relative to the MiniFS tree):
 *   ZeroResiduals(PolyGrid*)                src/Kernels/cal_flux.h:24
 *   InviscidFlux(PolyGrid*, limit, level)   src/Kernels/cal_flux.h:27
 *   InviscidFlux_TDTree(...)                src/Kernels/cal_flux.h:41
 *   ComputeTimeStep(PolyGrid*)              src/Kernels/compute_timestep.h:24
 *   ForwardLUSGS(PolyGrid*, level)          src/Kernels/forward_lusgs.h:23
 *   CalculateGradient(PolyGrid*)            src/Kernels/cal_gradient.h:23
 *   PassInterfaceData(PolyGrid*)            src/Kernels/load_grid.h:39
 *   LoadGrid / Decoupling / PreprocessGrid  src/Kernels/load_grid.h:27,43,33
 *   driver loop                             src/main.cpp:198-291
 *   solver globals                          src/Base/para_field_global.h:32-273
 *
 * Conventions.  Every function returns an int status (MFS_OK = 0, negative
 * on error) and never exits the process; mfs_last_error() returns the
 * message of the calling thread's last failure.  Arrays are plain host
 * pointers in the reference's layout: per-variable blocks (rho,u,v,w,p),
 * ghost cell of boundary face k at index nTCell+k, res as 5 blocks of
 * nTCell, DQ and gradients as 5 blocks of nTCell+nBFace.  Device work is
 * asynchronous on the context's own HIP stream; mfs_sync waits for it.
 * A context is used by one host thread at a time.
 */
#ifndef MFS_H
#define MFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFS_OK 0
#define MFS_ERR_ARG (-1)
#define MFS_ERR_HIP (-2)
#define MFS_ERR_IO (-3)
#define MFS_ERR_MESH (-4)
#define MFS_ERR_STATE (-5)
#define MFS_ERR_NOMEM (-6)

/* Precision of a context. */
#define MFS_FP64 0  /* plain FP64, bitwise equal to the reference's serial arithmetic */
#define MFS_MIXED 1 /* FP32 operator (flux, time step, LU-SGS), FP64 residual/update/norm */

/*
 * Solver parameters: the mutable globals of src/Base/para_field_global.cpp
 * (:37-221) gathered into one per-solve struct.  mfs_params_default fills
 * the reference defaults; mfs_params_derive recomputes the freestream and
 * limit values from mach/alpha/beita/re/T with the formulas of
 * para_field_global.cpp:150-209.
 */
typedef struct mfs_params {
    /* inputs */
    double mach, alpha, beita; /* alpha, beita in degrees */
    double re, T, gam, gascon;
    double cfl_start, cfl_end;
    int cfl_nstep;
    double epsa_r;      /* entropy-fix constant alf_l = epsa_r */
    double lhs_omga;    /* LU-SGS lhs ratio */
    double ratio_dtmax; /* dt_max/dt_min limit */
    /* limits of the state, para_field_global.cpp:200-208: rho in
     * [ratio_rhop_min rho_s, ratio_rhop_max rho_stag], absolute p in
     * [ratio_rhop_min p_bar, ratio_rhop_max (p_stag + p_bar)], CFL reduced
     * below absolute p = ratio_p_break p_bar (defaults 10, 1e-4, 0.01) */
    double ratio_rhop_max, ratio_rhop_min, ratio_p_break;
    /* derived by mfs_params_derive */
    double amu, ainf, uqq;
    double u_s, v_s, w_s, p_s, rho_s, p_bar, gamm1;
    double p_stag, e_stag, p_min, p_max, rho_min, rho_max, p_break, e_stag_max;
    double alf_l, alf_n;
} mfs_params;

void mfs_params_default(mfs_params* p);
int mfs_params_derive(mfs_params* p);

const char* mfs_last_error(void);

/* ------------------------------------------------------------------ */
/* Host grid: the PolyGrid of src/Base/grid_polyhedra.h after the set-up of
 * src/main.cpp:166-170.  Built and owned by the library.                */
typedef struct mfs_grid mfs_grid;

typedef struct mfs_grid_info {
    int nTCell, nTFace, nBFace, nTNode;
    int preprocessed; /* 1 after mfs_grid_preprocess */
    int decoupled;    /* 1 after mfs_grid_decouple */
} mfs_grid_info;

/* Box + Gaussian-bump Kuhn-tetrahedron mesh (unit box, nx*ny*nz hexes x 6
 * tets, interior nodes jittered by +-jitter*spacing, z=0 wall displaced by
 * bump_height*exp(-r^2/(2 bump_sigma^2)); patch 1 WALL, patch 2 FAR_FIELD). */
int mfs_grid_generate_box_bump(int nx, int ny, int nz, double jitter, double bump_height, double bump_sigma,
                               uint64_t seed, mfs_grid** out);
/* The same box with another cell kind: MFS_CELLS_TET (as above), MFS_CELLS_HEX
 * (one hexahedron per hex, 6 quadrilateral faces) or MFS_CELLS_PRISM (two
 * triangular prisms per hex).  The reference reads any polyhedral FlowStar
 * mesh (load_grid.cpp:129-273); these exercise its non-tetrahedral paths. */
#define MFS_CELLS_TET 0
#define MFS_CELLS_HEX 1
#define MFS_CELLS_PRISM 2
int mfs_grid_generate_box(int nx, int ny, int nz, int cells, double jitter, double bump_height, double bump_sigma,
                          uint64_t seed, mfs_grid** out);
/* LoadGrid with explicit paths (load_grid.cpp:82-314: FlowStar binary + bc.para). */
int mfs_grid_load(const char* mesh_path, const char* bc_path, mfs_grid** out);
int mfs_grid_save(const mfs_grid* g, const char* mesh_path, const char* bc_path);
/* Decoupling: TDTree decomposition and cell/node/face renumbering
 * (load_grid.cpp:339-395 under -DTDTREE).  Must precede preprocess. */
int mfs_grid_decouple(mfs_grid* g);
/* The same with the tree's parameters: nb_parts parts per tree node (the
 * reference's CreateTree uses 4, grid_polyhedra.cpp:378-383) and leaves
 * below part_size cells (512 there, :405).  The cell order sets the
 * LU-SGS dependency depth: on the 100^3 x 6 box+bump mesh 4/512 gives 342
 * sweep levels and 2/512 gives 129.  Any order is a valid input to the
 * path; results match the reference run on the same renumbered mesh.
 * Accepts 2 <= nb_parts <= 4 and part_size >= 8. */
int mfs_grid_decouple_ex(mfs_grid* g, int nb_parts, int part_size);
/* PreprocessGrid (load_grid.cpp:405-540): SpecifyBC, ComputeMetrics,
 * AllocAndInitFlowfield (freestream), SetGhostVariables. */
int mfs_grid_preprocess(mfs_grid* g, const mfs_params* p);
/* The same with ComputeMetrics (grid_polyhedra_auxiliary_functions.cpp:
 * 68-808: face centres, normals and areas, cell volumes and centroids,
 * ghost geometry, the normal and centroid corrections) on the current HIP
 * device; the cell-to-face lists are built there too.  The results land in
 * the grid's arrays, bitwise equal to mfs_grid_preprocess. */
int mfs_grid_preprocess_device(mfs_grid* g, const mfs_params* p);
int mfs_grid_get_info(const mfs_grid* g, mfs_grid_info* info);
/* Copies a named grid array ("x","f2c","xfn","area","vol","xcc","rho",
 * "bctype", "tdtree_leaves", ...) into dst; bytes must equal its size.
 * mfs_grid_array_bytes returns that size (0 if unknown). */
size_t mfs_grid_array_bytes(const mfs_grid* g, const char* name);
int mfs_grid_get_array(const mfs_grid* g, const char* name, void* dst, size_t bytes);
void mfs_grid_destroy(mfs_grid* g);

/* ------------------------------------------------------------------ */
/* Mesh view: borrowed host pointers to a preprocessed PolyGrid's arrays
 * (names as in grid_polyhedra.h).  The context copies what it needs.     */
typedef struct mfs_mesh_view {
    int nTCell, nTFace, nBFace, nTNode;
    const double *x, *y, *z;          /* nTNode */
    const int* nNPF;                  /* nTFace */
    const int* f2n;                   /* sum(nNPF) */
    const int* f2c;                   /* 2*nTFace, ghost of boundary face k = nTCell+k */
    const int* bc_type;               /* nBFace: WALL=3 SYMM=4 FAR_FIELD=6 ... */
    const double *xfn, *yfn, *zfn, *area; /* nTFace */
    const double *xfc, *yfc, *zfc;    /* nTFace */
    const double *xcc, *ycc, *zcc, *vol; /* nTCell+nBFace */
    const double *rho, *u, *v, *w, *p; /* nTCell+nBFace, p is gauge p - p_bar */
} mfs_mesh_view;

int mfs_grid_view(const mfs_grid* g, mfs_mesh_view* view);

/* ------------------------------------------------------------------ */
/* Device context: all device buffers and one HIP stream on the current
 * device.                                                               */
typedef struct mfs_ctx mfs_ctx;

int mfs_ctx_create(const mfs_mesh_view* mesh, const mfs_params* p, int precision, mfs_ctx** out);

/* Layout and launch options (no counterpart in the reference): results do
 * not depend on them, only speed.  Zero fields take the defaults. */
typedef struct mfs_ctx_options {
    int tile_cells;     /* cells per storage tile of the LU-SGS sweeps (>= 64; default 4096) */
    int sweep_waves;    /* waves per sweep workgroup: 4, 8 or 16 (default 4) */
    int sweep_by_tiles; /* 0: flat level-order schedule (default); 1: one workgroup per tile */
    int gradient_stream; /* mfs_step's CalculateGradient: 1 on the context stream, fused with the next
                          * residual pass; 2, 3, 4, 5 on a second stream of the context, started after the
                          * update, after the next iteration's residual pass, after its forward sweep, or
                          * after the wide levels of its forward sweep, when the thin tail begins
                          * (nothing in an iteration reads the gradient).  Default 5 (the fastest on 6.0M
                          * cells in both precisions). */
    int sweep_tail_grid; /* workgroups of the sweeps' thin-tail launch: each sweep of the flat schedule runs
                          * its wide levels and the run of thin levels that ends its order (at most 64
                          * chunks each) as two launches.  0: default (64); -1: one launch per sweep. */
} mfs_ctx_options;
int mfs_ctx_create_ex(const mfs_mesh_view* mesh, const mfs_params* p, int precision, const mfs_ctx_options* opt,
                      mfs_ctx** out);
int mfs_ctx_destroy(mfs_ctx* ctx);

/* The kernel entry points, one per reference function. */
int mfs_zero_residuals(mfs_ctx* ctx);     /* cal_flux.cpp:54-72 */
int mfs_inviscid_flux(mfs_ctx* ctx);      /* cal_flux.cpp:84-192, 1219-1254 (res -= / += F) */
int mfs_compute_timestep(mfs_ctx* ctx);   /* compute_timestep.cpp:66-122, 756-998, 1315-1338 */
int mfs_forward_lusgs(mfs_ctx* ctx);      /* forward_lusgs.cpp:67-163 incl. sweeps and update */
int mfs_pass_interface_data(mfs_ctx* ctx); /* load_grid.cpp:912-943: single domain, no-op */
/* SetGhostVariables (load_grid.cpp:953-1212, inviscid branch) on the device
 * state: the ghost cell of every boundary face from the cell in front of
 * it.  The reference calls it once during set-up and leaves the
 * per-iteration call commented out (main.cpp:283); mfs_ctx_set_ghost_refresh
 * makes mfs_step call it after each update, before the gradient (off by
 * default).  The FAR_FIELD subsonic branch uses pow, whose GPU result may
 * differ from the host's in the last bit. */
int mfs_set_ghost_variables(mfs_ctx* ctx);
int mfs_ctx_set_ghost_refresh(mfs_ctx* ctx, int enable);
int mfs_calculate_gradient(mfs_ctx* ctx); /* cal_gradient.cpp:87-167 */
/* iter_done++ of main.cpp:289 and one history record (norms of the last
 * residual, running dt_max/dt_min). */
int mfs_iteration_done(mfs_ctx* ctx);

/* n iterations of the driver loop (main.cpp:201-291), fused kernels;
 * appends n history records. */
int mfs_step(mfs_ctx* ctx, int n);
/* mfs_step for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops that forbid hipMalloc); it has no
 * counterpart in the reference.  It calls no hipMalloc, hipFree or other
 * allocating routine: the reduction partials, sweep ticket counters and
 * the block counters of the second stream's claiming kernels live in a
 * static device array of the library (2.03 MiB per GPU, held until the
 * process ends), and the n history records (7 doubles each, as
 * mfs_get_history gives them) go to hist_dev, a device buffer of 7*n
 * doubles that the caller owns.  Results are those of mfs_step(ctx, n);
 * the records are not added to the context's own history and kernel
 * timing is not recorded.  Limits: contexts of at most 8,388,608 cells
 * (MFS_ERR_ARG above that), not a batch member.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (contexts stepped
 * from several threads at once) race and must be serialised by the caller. */
int mfs_step_noalloc(mfs_ctx* ctx, int n, double* hist_dev);
int mfs_sync(mfs_ctx* ctx);

/* Per-iteration history: record k = {norm_rho, norm_mx, norm_my, norm_mz,
 * norm_et, dt_max, dt_min} (main.cpp:294-329 evaluated every iteration).
 * Copies min(max_records, available) records; returns the count (>=0). */
int mfs_history_count(mfs_ctx* ctx);
int mfs_get_history(mfs_ctx* ctx, double* out, int max_records);

/* Field ids for download/upload. */
#define MFS_F_RHO 0
#define MFS_F_U 1
#define MFS_F_V 2
#define MFS_F_W 3
#define MFS_F_P 4
#define MFS_F_RES 5  /* 5*nTCell */
#define MFS_F_DT 6   /* nTCell */
#define MFS_F_DQ 7   /* 5*(nTCell+nBFace) */
#define MFS_F_DQDX 8 /* 5*(nTCell+nBFace) */
#define MFS_F_DQDY 9
#define MFS_F_DQDZ 10
#define MFS_F_DIAG 11 /* nTCell, LU-SGS diagonal of the last ForwardLUSGS */
#define MFS_F_QNODE 12 /* 5*nTNode node values of the last gradient */
size_t mfs_field_count(mfs_ctx* ctx, int field);
int mfs_download(mfs_ctx* ctx, int field, double* dst, size_t count);
int mfs_upload(mfs_ctx* ctx, int field, const double* src, size_t count); /* state fields only */

/* Scalars of the running solve. */
typedef struct mfs_scalars {
    int iter_done;
    double dt_max, dt_min, cfl_min;
} mfs_scalars;
int mfs_get_scalars(mfs_ctx* ctx, mfs_scalars* s);

/* Timing and schedule facts for the bench and tests. */
typedef struct mfs_ctx_info {
    int nTCell, nTFace, nBFace, nTNode;
    int precision;
    int sweep_levels;      /* DAG depth of the LU-SGS sweep order */
    size_t device_bytes;   /* device memory held by the context */
    int tile_cells, sweep_waves, sweep_tiles, sweep_workgroups, sweep_by_tiles;
    int compute_units;     /* CUs of the context's device (the sweep grid is sized on them) */
    int sweep_tail_chunks; /* chunks of the sweeps' thin-tail launch (0: one launch per sweep) */
    int sweep_tail_cells;  /* their cells */
    int sweep_tail_workgroups;
} mfs_ctx_info;
int mfs_ctx_get_info(mfs_ctx* ctx, mfs_ctx_info* info);
/* Kernel time accounting: when enabled, every launch of the named kernel
 * group is bracketed with HIP events on the stream it runs on.  enable: 0
 * off, 1 every kernel group, 2 the two sweeps only.  The sweeps' events are
 * attached to their launches (they cost the stream nothing); the others
 * are event records, each of which leaves the stream idle for a few us
 * between two kernels. */
int mfs_ctx_enable_timing(mfs_ctx* ctx, int enable);
/* names: "flux","timestep","flux_ts" (both fused, mfs_step),"reduce","diag","clamp",
 * "fwd_sweep","bwd_sweep" (a sweep from its first launch's start to its last
 * launch's end), "fwd_sweep_wide","fwd_sweep_tail","bwd_sweep_tail",
 * "bwd_sweep_wide" (each launch of a sweep run as two, wide levels and thin
 * tail),"ghosts","node","gradient","grad_flux_ts" (mfs_step on tetrahedra: a
 * gradient with the next iteration's flux_ts);
 * returns accumulated milliseconds and launch count. */
int mfs_ctx_get_timing(mfs_ctx* ctx, const char* name, double* ms, int* launches);
int mfs_ctx_reset_timing(mfs_ctx* ctx);

/* ------------------------------------------------------------------ */
/* Batched multi-solve.  No counterpart in the reference, which runs one
 * solve per process (src/main.cpp:198-291).  A batch steps several
 * independent solves, each with its own mesh, flow condition and device
 * arrays, with one launch per kernel for all of them (the sweeps walk the
 * solves' chunks interleaved, each solve's in its own level order).  Each solve's arithmetic is the
 * single-solve path's, so its results are the same alone and in a batch.
 * All solves of a batch share one precision and one cell family
 * (tetrahedral meshes, or meshes with up to 6 faces per cell).          */
typedef struct mfs_batch mfs_batch;
int mfs_batch_create(int precision, mfs_batch** out);
/* Options of a batch (no counterpart in the reference); zero fields take
 * the defaults.  Results do not depend on them. */
typedef struct mfs_batch_options {
    int gradient_stream; /* the batch step's CalculateGradient: 1 on the batch stream, fused with the next
                          * residual pass; 2 or 4 on a second stream, started after the update or after the
                          * next iteration's forward sweep.  Default 1 for MFS_FP64, 4 for MFS_MIXED. */
} mfs_batch_options;
int mfs_batch_create_ex(int precision, const mfs_batch_options* opt, mfs_batch** out);
int mfs_batch_destroy(mfs_batch* b);
/* Builds the solve's device layout and uploads it; *slot names it until
 * removed.  mfs_batch_ctx(b, slot) gives its context for mfs_get_history,
 * mfs_download, mfs_get_scalars (do not destroy it directly). */
int mfs_batch_add(mfs_batch* b, const mfs_mesh_view* mesh, const mfs_params* p, int* slot);
/* Moves a context made by mfs_ctx_create(_ex), with the batch's precision
 * and device and in no batch yet, into the batch, which owns it from then
 * on (mfs_batch_add is create + attach).  Building a context is host work
 * proportional to the mesh; it may run on another thread while the batch
 * steps (the library keeps no shared mutable state), and attaching is
 * cheap.  A solve stepped alone before attaching continues where it stood.
 * On error the caller keeps the context. */
int mfs_batch_attach(mfs_batch* b, mfs_ctx* ctx, int* slot);
int mfs_batch_remove(mfs_batch* b, int slot);
int mfs_batch_step(mfs_batch* b, int n); /* n iterations of every solve of the batch */
int mfs_batch_sync(mfs_batch* b);
mfs_ctx* mfs_batch_ctx(mfs_batch* b, int slot);
int mfs_batch_count(mfs_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* MFS_H */
