// LU-SGS sweeps (SolveLUSGS3D, src/Kernels/forward_lusgs.cpp:803-1249) on
// the GPU, included by mfs_device.hip inside namespace mfs::dev.
//
// With LUSGSLayer[i] = i (ReorderCellforLUSGS_0, grid_polyhedra.cpp:1127-1142)
// the forward sweep of cell c reads the final DQ of its lower-numbered face
// neighbours and the backward sweep that of its higher-numbered neighbours
// and of its ghost cells (DQ = 0).
//
// Work decomposition: a wave handles one chunk of 64/LPC cells of the same
// forward DAG level (contiguous in device order), LPC lanes per cell, one
// lane per face slot.  Each lane evaluates FluxLUSGS3D for its own face;
// the lanes of a cell then exchange the products and every lane of the cell
// adds them in the cell's face order, so the sum is formed exactly as in
// the serial sweep.  The limiter and the update run redundantly on the
// lanes of a cell; each lane stores a share of the 5 results.
//
// Schedule: the forward sweep runs chunks in forward level order, the
// backward sweep in the reverse order.  Chunk u of a sweep is handled by
// wave u mod gridDim of a cooperative (co-resident) grid, each wave
// walking its chunks in increasing u: the lowest unfinished chunk is always
// being worked on and its dependencies are finished, so the sweep cannot
// deadlock.
//
// Publication: before a sweep every DQ slot it writes holds a NaN payload
// (kSentinel) that arithmetic never produces.  A cell writes its 5 final
// DQ values with device-coherent 64-bit stores; a dependent cell reads
// them with device-coherent loads and retries while any is the sentinel
// (each 64-bit store is single-copy atomic, so a non-sentinel value is
// final).  One lane per wave first waits on the chunk's latest dependency
// (rep), which keeps the polling traffic to one load per wave.  The forward
// sweep re-arms the backward array and the backward sweep the forward one.

constexpr unsigned long long kSentinel = 0x7FF4D51F5EED0000ULL;
constexpr int kMaxFaces = 6;  // tetrahedra, pyramids, prisms, hexahedra

__device__ __forceinline__ unsigned long long ld_dev(const double* p) {
    return __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<double*>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Value of lane S of this lane's group of LPC lanes (a quad for LPC = 4,
// by DPP; otherwise by a cross-lane permute).  Called with all lanes active.
template <int LPC, int S>
__device__ __forceinline__ double cell_bcast(double v) {
    if constexpr (LPC == 4) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, S * 0x55, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), S * 0x55, 0xf, 0xf, false);
        return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
    } else {
        return __shfl(v, (int)(threadIdx.x & ~(unsigned)(LPC - 1)) + S, 64);
    }
}

// True when another wave has already given up (checked every 4096 spins),
// so that one expired wait ends the whole sweep quickly.
__device__ __forceinline__ bool wait_failed(int* err, unsigned spins) {
    if ((spins & 4095u) != 0) return false;
    return __shfl(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0, 64) != 0;
}

struct SweepArgs {
    int nC, nchunks;
    const int2* chunk;   // forward order: first device cell, cell count
    const int* cfs;      // LPC face slots per device cell (-1 = none)
    const int2* fcell;
    const double4* fgeo;
    const double* q;     // state read by both sweeps (cell-major, 5 per cell)
    const double* res;   // forward: initial DQ
    const double* diag;
    double* dqf;         // forward DQ, 5 per cell
    double* dqb;         // backward (final) DQ, 5 per cell incl. ghosts (0)
    double* qout;        // backward: updated state
    const int* rep;      // per chunk in run order: dependency published last, -1 if none
    const int* depmask;  // per cell: forward / backward dependency face slots
    int cell0;           // cell with mesh id 0 (forward: divide only)
    int* err;
};

// Adds the LPC per-face products of this lane's cell in slot order (the
// cell's ascending face order), taking only the slots in `use`.
template <bool FWD, int LPC>
__device__ __forceinline__ void add_cell_terms(const double (&p)[5], unsigned use, double (&D)[5],
                                               double (&acc)[5]) {
#define MFS_ADD_SLOT(S)                                                   \
    if constexpr ((S) < LPC) {                                            \
        double ps[5];                                                     \
        _Pragma("unroll") for (int j = 0; j < 5; ++j) ps[j] = cell_bcast<LPC, (S)>(p[j]); \
        if (use & (1u << (S))) {                                          \
            _Pragma("unroll") for (int j = 0; j < 5; ++j) {               \
                if (FWD)                                                  \
                    D[j] -= ps[j];                                        \
                else                                                      \
                    acc[j] += ps[j];                                      \
            }                                                             \
        }                                                                 \
    }
    MFS_ADD_SLOT(0)
    MFS_ADD_SLOT(1)
    MFS_ADD_SLOT(2)
    MFS_ADD_SLOT(3)
    MFS_ADD_SLOT(4)
    MFS_ADD_SLOT(5)
    MFS_ADD_SLOT(6)
    MFS_ADD_SLOT(7)
#undef MFS_ADD_SLOT
}

template <bool FWD, int LPC>
__global__ __launch_bounds__(64) void k_sweep(SweepArgs A, FlowConst K) {
    constexpr int CPC = 64 / LPC;
    const int lane = (int)threadIdx.x;
    const int k = lane / LPC;  // cell of the chunk
    const int s = lane % LPC;  // face slot
    double* const poll = FWD ? A.dqf : A.dqb;
    // a sweep that follows a failed one in the same stream does not start
    if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    for (int u = blockIdx.x; u < A.nchunks; u += gridDim.x) {
        const int t = FWD ? u : A.nchunks - 1 - u;
        const int2 ch = A.chunk[t];
        const bool active = k < ch.y;
        const int c = ch.x + (active ? k : 0);  // idle lanes read the first cell and store nothing
        const unsigned m = (unsigned)A.depmask[c];
        unsigned kind = FWD ? (m & 0xffu) : ((m >> 8) & 0xffu);  // slots that take part
        const unsigned pendm = FWD ? kind : ((m >> 16) & 0xffu);  // slots with a DQ to wait for
        if (FWD && c == A.cell0) kind = 0;  // forward_lusgs.cpp:803-804: the first cell only divides
        const bool mine = active && ((kind >> s) & 1u);
        bool pend = active && ((pendm >> s) & 1u) && mine;
        // operands that do not depend on the sweep, loaded before waiting
        double fnx = 0.0, fny = 0.0, fnz = 0.0, far_ = 0.0;
        double qnb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double dv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        int nb = 0;
        if (mine) {
            const int f = A.cfs[(size_t)c * LPC + s];
            const int2 cc = A.fcell[f];
            const double4 g = A.fgeo[f];
            const bool flip = (cc.y == c);  // normal oriented from c to its neighbour
            nb = flip ? cc.x : cc.y;
            fnx = flip ? -g.x : g.x;
            fny = flip ? -g.y : g.y;
            fnz = flip ? -g.z : g.z;
            far_ = g.w;
            load5(A.q + (size_t)nb * 5, qnb);
        }
        double qc[5], D[5];
        load5(A.q + (size_t)c * 5, qc);
        const double dg = A.diag[c];
        load5((FWD ? A.res : A.dqf) + (size_t)c * 5, D);

        unsigned spins = 0;
        // one lane waits for the chunk's latest dependency
        const int rep = A.rep[u];
        if (rep >= 0) {
            while (true) {
                unsigned long long v = 0;
                if (lane == 0) v = ld_dev(poll + (size_t)rep * 5);
                v = __shfl(v, 0, 64);
                if (v != kSentinel) break;
                if (++spins > kSpinLimit || wait_failed(A.err, spins)) {
                    if (lane == 0) atomicExch(A.err, 1);
                    return;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        // every lane reads its face's dependency, retrying the rare unpublished one
        while (true) {
            if (pend) {
                unsigned long long b[5];
                bool ok = true;
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    b[j] = ld_dev(poll + (size_t)nb * 5 + j);
                    ok = ok && (b[j] != kSentinel);
                }
                if (ok) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) dv[j] = __builtin_bit_cast(double, b[j]);
                    pend = false;
                }
            }
            if (__ballot(pend) == 0) break;
            if (++spins > kSpinLimit || wait_failed(A.err, spins)) {
                if (lane == 0) atomicExch(A.err, 1);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // this lane's face term: forward 0.5*A*F (DQ -= term), backward A*F (summed)
        double p[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (mine) {
            double fl[5];
            flux_lusgs<double>(qnb, dv, fnx, fny, fnz, K.gam, K.p_bar, K.lhs_omga, fl);
            const double tmp = FWD ? 0.5 * far_ : far_;
#pragma unroll
            for (int j = 0; j < 5; ++j) p[j] = tmp * fl[j];
        }
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        add_cell_terms<FWD, LPC>(p, kind, D, acc);
        if (FWD) {
#pragma unroll
            for (int j = 0; j < 5; ++j) D[j] /= dg;
            if (c != A.cell0) dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
        } else {
            const double tmp = 2.0 * dg;
#pragma unroll
            for (int j = 0; j < 5; ++j) D[j] -= acc[j] / tmp;
            dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
        }
        if (active) {
            double* const own = FWD ? A.dqf : A.dqb;    // published
            double* const other = FWD ? A.dqb : A.dqf;  // re-armed for the next sweep
            const double sent = __builtin_bit_cast(double, kSentinel);
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                if (j % LPC == s) {
                    st_dev(own + (size_t)c * 5 + j, D[j]);
                    other[(size_t)c * 5 + j] = sent;
                }
            }
        }
        if (!FWD) {
            double qnew[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) qnew[j] = qc[j];
            update_cell(qnew, D, K);
            if (active) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    if (j % LPC == s) A.qout[(size_t)c * 5 + j] = qnew[j];
            }
        }
        (void)CPC;
    }
}
