// Batched multi-solve kernels, included by mfs_device.hip inside namespace
// mfs::dev after the per-solve kernels.
//
// A batch steps several independent solves (each its own mesh, flow
// condition and device arrays) with one launch per kernel for all of them.
// Each launch reads a table with one BatchSolve entry per solve; the cell
// and node kernels map their workgroups to solves through a block map (each
// solve owns a whole number of XCD groups of 8 workgroups).  The per-cell arithmetic is the
// single-solve kernels' own (the same device functions), so a solve gives
// the same bits alone and in a batch.  The sweeps walk a list of (solve, u)
// items that interleaves the solves' chunks, handed out by ticket groups.
// The forward list aligns the solves at their last chunk and the backward
// list at their first, so that the thin tails of all solves' level orders
// run side by side.  Dependencies never cross solves and each solve's
// chunks keep their order, so the deadlock argument of the single-solve
// sweep holds.

template <typename R>
struct BatchSolve {
    int nC, nN, n;                // cells, nodes, cells + ghosts
    int blk0_cell, nblk_cell;     // global block range of the cell kernels
    int blk0_node, nblk_node;     // and of the node kernel
    const int *cfo, *cf;
    const int2* fcell;
    const double4* fgeo;
    const double2* fdn;
    const double* vol;
    const double* q;              // state at the start of the iteration
    const double* qnew;           // state written by the backward sweep
    double *res, *dt, *diag, *part, *run, *hist;
    int hist_slot;
    double cfl, cfl_min, ratio_dtmax;
    float *q32, *diag32;          // mixed precision only
    FlowConst K;
    // gradient
    const int* nlo;
    const int* ell_off;
    const int2* ell_ab;
    const R* ell_w;
    const double* nden;
    R* qn;
    const int *fno, *fn, *btype;
    const typename Real<R>::vec4* fgeo_r;
    R *gx, *gy, *gz;
    const int* snode;             // tetrahedral meshes: slot nodes (nullptr otherwise)
    const double4* sgeo64;        // FP64 slot geometry, distances and sides (residual kernel)
    const double* sdn;
    const int* sflip;
    // sweeps
    SweepArgs<R> sw;
};

template <int MAXF, bool MIXED, typename R>
__global__ __launch_bounds__(kBlock) void k_b_flux_ts(const BatchSolve<R>* __restrict__ T,
                                                      const int* __restrict__ cell_solve) {
    const BatchSolve<R>& S = T[cell_solve[blockIdx.x]];
    const int lb = (int)blockIdx.x - S.blk0_cell;
    const int lg = (lb % kXcd) * (S.nblk_cell / kXcd) + lb / kXcd;  // XCD-aware inside the solve
    flux_ts_slot_body<MAXF == 4 ? 4 : 8, MIXED>(lg * kBlock + (int)threadIdx.x, S.nC, S.sw.snb, S.sgeo64, S.sdn,
                                                S.sflip, S.q, S.vol, S.res, S.dt, S.diag, S.part + lb, S.nblk_cell, S.K,
                                                S.cfl, S.cfl_min, S.q32, S.diag32);
}

// residual norms, dt extrema and the history record of every solve (one
// workgroup per solve)
template <typename R>
__global__ __launch_bounds__(kReduceBlock) void k_b_reduce(const BatchSolve<R>* __restrict__ T) {
    const BatchSolve<R>& S = T[blockIdx.x];
    reduce_body(S.nblk_cell, S.part, S.run, S.hist, S.hist_slot, 1 | 2 | 4, S.ratio_dtmax);
}

template <int MAXF, typename R>
__global__ __launch_bounds__(kBlock) void k_b_clamp(const BatchSolve<R>* __restrict__ T,
                                                    const int* __restrict__ cell_solve) {
    const BatchSolve<R>& S = T[cell_solve[blockIdx.x]];
    const int lb = (int)blockIdx.x - S.blk0_cell;
    const int lg = (lb % kXcd) * (S.nblk_cell / kXcd) + lb / kXcd;
    diag_body<MAXF>(lg * kBlock + (int)threadIdx.x, S.nC, S.cfo, S.cf, S.fgeo, S.q, S.vol, S.dt, S.diag, S.run,
                    S.ratio_dtmax, S.K, S.q32 ? S.diag32 : nullptr);
}

template <bool FWD, int LPC, typename R>
__global__ __launch_bounds__(kSweepBlock) void k_b_sweep(const BatchSolve<R>* __restrict__ T,
                                                        const int2* __restrict__ items, int nitems, int* tickets,
                                                        int* tickets_next, int ngroups) {
    // items: (solve, u) with u the position in that solve's level order.
    // Item i belongs to ticket group i mod ngroups; the waves of the
    // workgroups of group g take its items in order from the group's
    // counter, as in the single-solve sweep.
    if (blockIdx.x == 0 && threadIdx.x < kTicketGroups) tickets_next[threadIdx.x * kTicketStride] = 0;
    const int grp = (int)blockIdx.x % ngroups;
    int* const ctr = tickets + grp * kTicketStride;
    const int lane = (int)threadIdx.x & 63;
    int tk = 0;
    if (lane == 0) tk = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tk = __shfl(tk, 0, 64);
    while (true) {
        const int it = grp + ngroups * tk;
        if (it >= nitems) return;
        int next = 0;
        if (lane == 0) next = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int2 su = items[it];
        const SweepArgs<R>& A = T[su.x].sw;
        if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            (void)sweep_chunk<FWD, LPC, R, false>(A, T[su.x].K, FWD ? su.y : A.nchunks - 1 - su.y);
        tk = __shfl(next, 0, 64);
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_b_node(const BatchSolve<R>* __restrict__ T,
                                                   const int* __restrict__ node_solve) {
    const BatchSolve<R>& S = T[node_solve[blockIdx.x]];
    const int lb = (int)blockIdx.x - S.blk0_node;
    const int lg = (lb % kXcd) * (S.nblk_node / kXcd) + lb / kXcd;
    node_body<R>(lg * kBlock + (int)threadIdx.x, lg * (kBlock / 64) + (int)threadIdx.x / 64, S.nN, S.nlo, S.ell_off,
                 S.ell_ab, S.ell_w, S.nden, S.qnew, S.qn);
}

template <int MAXF, typename R>
__global__ __launch_bounds__(kBlock) void k_b_grad(const BatchSolve<R>* __restrict__ T,
                                                   const int* __restrict__ cell_solve) {
    const BatchSolve<R>& S = T[cell_solve[blockIdx.x]];
    const int lb = (int)blockIdx.x - S.blk0_cell;
    const int lg = (lb % kXcd) * (S.nblk_cell / kXcd) + lb / kXcd;
    if (MAXF == 4 && S.snode)
        grad_tet_body<R>(lg * kBlock + (int)threadIdx.x, S.nC, S.n, S.sw.snb, S.sw.sgeo, S.snode, S.sw.depmask,
                         S.qnew, S.qn, S.vol, S.gx, S.gy, S.gz);
    else
        grad_body<MAXF, R>(lg * kBlock + (int)threadIdx.x, S.nC, S.n, S.cfo, S.cf, S.fcell, S.fgeo_r, S.fno, S.fn,
                           S.btype, S.qnew, S.qn, S.vol, S.gx, S.gy, S.gz);
}

// The node values and the gradient of the batch on a small grid whose
// workgroups claim blocks (on the second stream, beside the backward
// sweep), as k_grad_tet_claim does for one solve's gradient: workgroup
// b takes the next unclaimed global block b mod 8 + 8k from counter
// b mod 8, so an XCD walks each solve's blocks of that XCD in order.  The
// block counts are multiples of 8 (each solve owns whole groups of 8); the
// counters are 64 ints apart, zero before the launch and left at zero.  64
// seeded solves, mixed precision, one workgroup per CU: 11.52 -> 11.34 ms
// per batch iteration (two per CU: no gain; fixed shares per workgroup:
// 11.46).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_b_node_claim(const BatchSolve<R>* __restrict__ T,
                                                         const int* __restrict__ node_solve, int nblk, int* counters) {
    __shared__ int claimed;
    const int x = (int)blockIdx.x % kXcd, per = (int)gridDim.x / kXcd, span = nblk / kXcd;
    int* const ctr = counters + x * 64;
    if (threadIdx.x == 0) claimed = atomicAdd(ctr, 1);
    __syncthreads();
    int k = claimed;
    while (k < span) {
        __syncthreads();  // every thread has read `claimed`
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1);
        const int gb = k * kXcd + x;
        const BatchSolve<R>& S = T[node_solve[gb]];
        const int lb = gb - S.blk0_node;
        const int lg = (lb % kXcd) * (S.nblk_node / kXcd) + lb / kXcd;
        node_body<R>(lg * kBlock + (int)threadIdx.x, lg * (kBlock / 64) + (int)threadIdx.x / 64, S.nN, S.nlo, S.ell_off,
                     S.ell_ab, S.ell_w, S.nden, S.qnew, S.qn);
        if (threadIdx.x == 0) claimed = next;
        __syncthreads();
        k = claimed;
    }
    if (threadIdx.x == 0 && k == span + per - 1) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int MAXF, typename R>
__global__ __launch_bounds__(kBlock) void k_b_grad_claim(const BatchSolve<R>* __restrict__ T,
                                                         const int* __restrict__ cell_solve, int nblk, int* counters) {
    __shared__ int claimed;
    const int x = (int)blockIdx.x % kXcd, per = (int)gridDim.x / kXcd, span = nblk / kXcd;
    int* const ctr = counters + x * 64;
    if (threadIdx.x == 0) claimed = atomicAdd(ctr, 1);
    __syncthreads();
    int k = claimed;
    while (k < span) {
        __syncthreads();  // every thread has read `claimed`
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1);
        const int gb = k * kXcd + x;
        const BatchSolve<R>& S = T[cell_solve[gb]];
        const int lb = gb - S.blk0_cell;
        const int lg = (lb % kXcd) * (S.nblk_cell / kXcd) + lb / kXcd;
        if (MAXF == 4 && S.snode)
            grad_tet_body<R>(lg * kBlock + (int)threadIdx.x, S.nC, S.n, S.sw.snb, S.sw.sgeo, S.snode, S.sw.depmask,
                             S.qnew, S.qn, S.vol, S.gx, S.gy, S.gz);
        else
            grad_body<MAXF, R>(lg * kBlock + (int)threadIdx.x, S.nC, S.n, S.cfo, S.cf, S.fcell, S.fgeo_r, S.fno, S.fn,
                               S.btype, S.qnew, S.qn, S.vol, S.gx, S.gy, S.gz);
        if (threadIdx.x == 0) claimed = next;
        __syncthreads();
        k = claimed;
    }
    if (threadIdx.x == 0 && k == span + per - 1) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The gradient of iteration i (table T) fused with the residual / time-step
// pass of iteration i+1 (table T1), as in the single-solve step: the same
// solves in both tables, the gradient feeding nothing of iteration i+1.
template <int MAXF, bool MIXED, typename R>
__global__ __launch_bounds__(kBlock) void k_b_grad_flux_ts(const BatchSolve<R>* __restrict__ T,
                                                          const BatchSolve<R>* __restrict__ T1,
                                                          const int* __restrict__ cell_solve) {
    const int a = cell_solve[blockIdx.x];
    const BatchSolve<R>& S = T[a];
    const BatchSolve<R>& S1 = T1[a];
    const int lb = (int)blockIdx.x - S.blk0_cell;
    const int lg = (lb % kXcd) * (S.nblk_cell / kXcd) + lb / kXcd;
    const int c = lg * kBlock + (int)threadIdx.x;
    if (MAXF == 4 && S.snode)
        grad_tet_body<R>(c, S.nC, S.n, S.sw.snb, S.sw.sgeo, S.snode, S.sw.depmask, S.qnew, S.qn, S.vol, S.gx, S.gy,
                         S.gz);
    else
        grad_body<MAXF, R>(c, S.nC, S.n, S.cfo, S.cf, S.fcell, S.fgeo_r, S.fno, S.fn, S.btype, S.qnew, S.qn, S.vol,
                           S.gx, S.gy, S.gz);
    flux_ts_slot_body<MAXF == 4 ? 4 : 8, MIXED>(c, S1.nC, S1.sw.snb, S1.sgeo64, S1.sdn, S1.sflip, S1.q, S1.vol,
                                                S1.res, S1.dt, S1.diag, S1.part + lb, S1.nblk_cell, S1.K, S1.cfl,
                                                S1.cfl_min, S1.q32, S1.diag32);
}
