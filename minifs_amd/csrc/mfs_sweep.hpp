// LU-SGS sweeps (SolveLUSGS3D, src/Kernels/forward_lusgs.cpp:803-1249) on
// the GPU, included by mfs_device.hip inside namespace mfs::dev.
//
// With LUSGSLayer[i] = i (ReorderCellforLUSGS_0, grid_polyhedra.cpp:1127-1142)
// the forward sweep of cell c reads the final DQ of its lower-numbered face
// neighbours and the backward sweep that of its higher-numbered neighbours
// and of its ghost cells (DQ = 0).
//
// Schedule: cells sorted by DAG level, each level padded to whole 64-cell
// chunks, so the 64 cells of a chunk never depend on one another and all
// their dependencies lie in earlier chunks.  Chunk t is handled by wave
// t mod gridDim of a cooperative (co-resident) grid, each wave walking its
// chunks in increasing order: the lowest unfinished chunk is always being
// worked on and its dependencies are finished, so the sweep cannot
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
//
// The per-cell arithmetic is the serial sweep's, in the serial face order,
// so DQ and the updated state are bitwise equal to the reference's.

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

struct SweepArgs {
    int nC, nchunks;
    const int* cfo;
    const int* cf;
    const int2* fcell;
    const double4* fgeo;
    const double* q;     // state read by both sweeps (cell-major, 5 per cell)
    const double* res;   // forward: initial DQ
    const double* diag;
    double* dqf;         // forward DQ, 5 per cell
    double* dqb;         // backward (final) DQ, 5 per cell incl. ghosts (0)
    double* qout;        // backward: updated state
    const int* sched;    // padded level schedule (-1 = empty slot)
    const int* rep;      // per chunk: dependency published last, -1 if none
    const int* depmask;  // per cell: forward / backward dependency face slots
    int cell0;           // cell with mesh id 0 (forward: divide only)
    int* err;
};

template <bool FWD, int MAXF>
__global__ __launch_bounds__(64) void k_sweep(SweepArgs A, FlowConst K) {
    const int nC = A.nC;
    double* const poll = FWD ? A.dqf : A.dqb;
    for (int t = blockIdx.x; t < A.nchunks; t += gridDim.x) {
        const int c = A.sched[t * 64 + (int)threadIdx.x];
        const bool active = c >= 0;
        int nf = 0;
        int fs[MAXF], nbs[MAXF];
        unsigned kind = 0;  // bit s: face s takes part in this sweep
        unsigned pend = 0;  // bit s: dependency whose DQ has not been read yet
        double qc[5], D[5], dg = 1.0;
        double dv[MAXF][5];
        // operands that do not depend on the sweep, loaded before waiting:
        // the face normal oriented from c to its neighbour, the face area
        // and the neighbour's state
        double fnx[MAXF], fny[MAXF], fnz[MAXF], far_[MAXF], qnb[MAXF][5];
        if (active) {
            nf = cell_faces<MAXF>(A.cfo, A.cf, c, fs);
            // dependency slots, decided on mesh ids when the layout was built
            const unsigned m = (unsigned)A.depmask[c];
            kind = FWD ? (m & 0xffu) : ((m >> 8) & 0xffu);
            pend = FWD ? kind : ((m >> 16) & 0xffu);  // ghost cells have DQ = 0
#pragma unroll
            for (int s = 0; s < MAXF; ++s) {
                const int2 cc = A.fcell[fs[s]];
                const int nb = (cc.x == c) ? cc.y : cc.x;
                nbs[s] = nb;
#pragma unroll
                for (int j = 0; j < 5; ++j) dv[s][j] = 0.0;
                const double4 g = A.fgeo[fs[s]];
                const bool flip = (cc.y == c);
                fnx[s] = flip ? -g.x : g.x;
                fny[s] = flip ? -g.y : g.y;
                fnz[s] = flip ? -g.z : g.z;
                far_[s] = g.w;
                load5(A.q + (size_t)nb * 5, qnb[s]);
            }
            load5(A.q + (size_t)c * 5, qc);
            dg = A.diag[c];
            if (FWD) {
                load5(A.res + (size_t)c * 5, D);
#pragma unroll
                for (int j = 0; j < 5; ++j) A.dqb[(size_t)c * 5 + j] = __builtin_bit_cast(double, kSentinel);
            } else {
                load5(A.dqf + (size_t)c * 5, D);
#pragma unroll
                for (int j = 0; j < 5; ++j) A.dqf[(size_t)c * 5 + j] = __builtin_bit_cast(double, kSentinel);
            }
        }
        unsigned spins = 0;
        // one lane waits for the chunk's latest dependency
        const int rep = A.rep[t];
        if (rep >= 0) {
            while (true) {
                unsigned long long v = 0;
                if (threadIdx.x == 0) v = ld_dev(poll + (size_t)rep * 5);
                v = __shfl(v, 0, 64);
                if (v != kSentinel) break;
                if (++spins > kSpinLimit) {
                    if (threadIdx.x == 0) atomicExch(A.err, 1);
                    return;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        // every lane reads its dependencies, retrying the rare unpublished one
        while (true) {
            if (pend) {
#pragma unroll
                for (int s = 0; s < MAXF; ++s) {
                    if (pend & (1u << s)) {
                        unsigned long long b[5];
                        bool ok = true;
#pragma unroll
                        for (int j = 0; j < 5; ++j) {
                            b[j] = ld_dev(poll + (size_t)nbs[s] * 5 + j);
                            ok = ok && (b[j] != kSentinel);
                        }
                        if (ok) {
#pragma unroll
                            for (int j = 0; j < 5; ++j) dv[s][j] = __builtin_bit_cast(double, b[j]);
                            pend &= ~(1u << s);
                        }
                    }
                }
            }
            if (__ballot(pend != 0) == 0) break;
            if (++spins > kSpinLimit) {
                if (threadIdx.x == 0) atomicExch(A.err, 1);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!active) continue;
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (!(FWD && c == A.cell0)) {  // forward_lusgs.cpp:803-804: the first cell only divides
#pragma unroll
            for (int s = 0; s < MAXF; ++s) {
                if (kind & (1u << s)) {
                    double fl[5];
                    flux_lusgs<double>(qnb[s], dv[s], fnx[s], fny[s], fnz[s], K.gam, K.p_bar, K.lhs_omga, fl);
                    if (FWD) {
                        const double tmp = 0.5 * far_[s];
#pragma unroll
                        for (int j = 0; j < 5; ++j) D[j] -= tmp * fl[j];
                    } else {
                        const double tmp = far_[s];
#pragma unroll
                        for (int j = 0; j < 5; ++j) acc[j] += tmp * fl[j];
                    }
                }
            }
        }
        if (FWD) {
#pragma unroll
            for (int j = 0; j < 5; ++j) D[j] /= dg;
            if (c != A.cell0) dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
#pragma unroll
            for (int j = 0; j < 5; ++j) st_dev(A.dqf + (size_t)c * 5 + j, D[j]);
        } else {
            const double tmp = 2.0 * dg;
#pragma unroll
            for (int j = 0; j < 5; ++j) D[j] -= acc[j] / tmp;
            dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
#pragma unroll
            for (int j = 0; j < 5; ++j) st_dev(A.dqb + (size_t)c * 5 + j, D[j]);
            double qnew[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) qnew[j] = qc[j];
            update_cell(qnew, D, K);
#pragma unroll
            for (int j = 0; j < 5; ++j) A.qout[(size_t)c * 5 + j] = qnew[j];
        }
    }
}
