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
// Schedule: cells are stored by tile (4096 consecutive mesh ids, a compact
// region in TDTree numbering), then forward level; a chunk is up to 64/LPC
// cells of one (tile, level) group, so its cells never depend on one
// another.  Two schedules, both on a grid sized to be resident at once:
//   flat (default): the chunks in forward level order (`run`), the backward
//     sweep walking it from the end; the steps go to the ticket groups in
//     runs of 2^run_shift consecutive steps, run k to group k mod ngroups,
//     and a free wave of the group takes the group's next step from the
//     group's counter, so waves that finish early take more chunks and no
//     wave's delay holds back the chunks after it.  A group's workgroups
//     share an XCD, so a run's chunks (neighbours in one tile and level,
//     whose rows share cache lines) meet in one L2.
//   tiles: a workgroup of W waves owns a tile, wave w taking the tile's
//     chunks w, w+W, ... in level order; tiles go to workgroups
//     round-robin, each workgroup walking its tiles in sweep order.
// In both, every dependency of a chunk comes earlier in the sweep order and
// each wave takes its chunks in that order, so the first unfinished chunk
// is always being worked on with its dependencies finished, or is the next
// ticket of a group whose waves are all free: the sweep cannot deadlock
// while every wave of the grid is resident.  (On 6M cells the flat
// schedule measured 2x faster: dependencies inside a tile cost the same
// device-coherent hop as across tiles, and a workgroup held by one tile
// delays its next.)
//
// Publication: before a sweep every DQ slot it writes holds a NaN payload
// (the sentinel) that arithmetic never produces.  A cell writes its 5 final
// DQ values with device-coherent stores; a dependent cell reads them with
// device-coherent loads and retries while any is the sentinel (each store
// is single-copy atomic, so a non-sentinel value is final).  One lane per
// wave first waits on the chunk's latest dependency (rep), which keeps the
// polling traffic to one load per wave; on thin levels (few chunks, few
// waves polling) every lane polls its own dependency at once, which saves
// the second round trip on the level's critical path.  The forward sweep re-arms the
// backward array and the backward sweep the forward one.
//
// Precision: R is the type of the implicit operator (state snapshot,
// geometry, Diag, DQ).  With R = double the sweeps are the reference's
// arithmetic bit for bit.  With R = float (mixed precision) the operator
// runs in FP32 while the right-hand side (the FP64 residual) and the state
// update stay FP64, so the iteration is a defect correction whose fixed
// point is the FP64 solution.

constexpr int kMaxFaces = 6;  // tetrahedra, pyramids, prisms, hexahedra

constexpr unsigned long long kSentinel = Real<double>::kSent;

template <typename R>
__device__ __forceinline__ typename Real<R>::bits ld_dev(const R* p) {
    using B = typename Real<R>::bits;
    return __hip_atomic_load(reinterpret_cast<B*>(const_cast<R*>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename R>
__device__ __forceinline__ void st_dev(R* p, R v) {
    using B = typename Real<R>::bits;
    __hip_atomic_store(reinterpret_cast<B*>(p), __builtin_bit_cast(B, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
template <int LPC, int S>
__device__ __forceinline__ float cell_bcast(float v) {
    if constexpr (LPC == 4) {
        const int r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), S * 0x55, 0xf, 0xf, false);
        return __builtin_bit_cast(float, r);
    } else {
        return __shfl(v, (int)(threadIdx.x & ~(unsigned)(LPC - 1)) + S, 64);
    }
}

// 5-value records read with 128-bit buffer loads: 3 loads for a 40-byte
// record, 2 for a 20-byte one, where per-value loads take 5.  The sweeps
// read their dependencies' published DQ this way: those loads bypass the
// CU's cache, so each of them is a request to the memory system (the
// plain loads of a chunk's own and its neighbours' records hit the CU's
// cache after the first and gain nothing from it).  Buffer offsets are
// 32-bit, so the sweeps take this path only while the state arrays are
// below 2 GiB (wide_records).
constexpr int kBufferDword3 = 0x00020000;  // raw buffer resource: byte-addressed, no format
constexpr int kLoadDevice = 16;            // the cache policy of ld_dev (device-coherent, sc1)
__device__ __forceinline__ bool wide_records(int nq) { return (size_t)nq * 5 * sizeof(double) <= 0x7fffffffu; }

template <int AUX, typename T>
__device__ __forceinline__ void ld5w(const T* base, int nrec, int i, T (&v)[5]) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), 0, nrec * 5 * (int)sizeof(T), kBufferDword3);
    const int off = i * 5 * (int)sizeof(T);
    if constexpr (sizeof(T) == 8) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, AUX);
        const auto c = __builtin_amdgcn_raw_buffer_load_b64(r, off + 32, 0, AUX);
        const unsigned w[10] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], c[0], c[1]};
#pragma unroll
        for (int j = 0; j < 5; ++j)
            v[j] = __builtin_bit_cast(T, ((unsigned long long)w[2 * j + 1] << 32) | w[2 * j]);
    } else {
        // the words go through plain integers: a bit_cast straight from an
        // element of the returned vector compiled to element 0 repeated
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        const unsigned b = __builtin_amdgcn_raw_buffer_load_b32(r, off + 16, 0, AUX);
        const unsigned w[5] = {a[0], a[1], a[2], a[3], b};
#pragma unroll
        for (int j = 0; j < 5; ++j) v[j] = __builtin_bit_cast(T, w[j]);
    }
}
// Record i of `base` (nrec records) as published by another wave, read
// once.  The wide path is not an atomic: single reads only, never in a
// wait loop (the compiler could keep its value across the loop's turns).
template <typename T>
__device__ __forceinline__ void read_published(bool wide, const T* base, int nrec, int i, T (&v)[5]) {
    if (wide) {
        ld5w<kLoadDevice>(base, nrec, i, v);
    } else {
#pragma unroll
        for (int j = 0; j < 5; ++j) v[j] = __builtin_bit_cast(T, ld_dev<T>(base + (size_t)i * 5 + j));
    }
}

// True when another wave has already given up (checked every 4096 spins),
// so that one expired wait ends the whole sweep quickly.
__device__ __forceinline__ bool wait_failed(int* err, unsigned spins) {
    if ((spins & 4095u) != 0) return false;
    return __shfl(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0, 64) != 0;
}

template <typename R>
struct SweepArgs {
    int nC, nchunks;
    int nq;                                 // records of the state arrays: interior and ghost cells
    int ntiles;
    const int* tile_chunk;                  // ntiles+1: first chunk of each tile (device order)
    const int* pos;                         // per chunk: its position in the flat order
    int flat;                               // 1: flat schedule, 0: tiles
    const int4* crun;                       // per position: first device cell, cell count, rep fwd, rep bwd
    const int* snb;                         // LPC per device cell: neighbour across the slot's face
    const typename Real<R>::vec4* sgeo;     // LPC per device cell: normal away from the cell + area
    const R* q;                             // operator state (cell-major, 5 per cell)
    const double* q64;                      // FP64 state the backward sweep updates
    const double* res;                      // forward: initial DQ (FP64 residual)
    const R* diag;
    R* dqf;                                 // forward DQ, 5 per cell
    R* dqb;                                 // backward (final) DQ, 5 per cell incl. ghosts (0)
    double* qout;                           // backward: updated state
    const int* depmask;                     // per cell: forward / backward dependency face slots
    int cell0;                              // cell with mesh id 0 (forward: divide only)
    int* err;
    int* tickets;                           // this sweep's ticket counters (flat schedule)
    int* tickets_next;                      // the other direction's, zeroed here for its next sweep
    int ngroups;                            // ticket groups in use (<= kTicketGroups and <= workgroups)
    // flat schedule: this launch walks sweep steps [ubeg, uend) of the
    // sweep's order (step u is position u forward, nchunks-1-u backward);
    // a sweep split in two launches runs its wide levels and its thin tail
    // apart, each launch ending when all its steps are done
    int ubeg, uend;
    // flat schedule: the steps go to the ticket groups in runs of
    // 2^run_shift consecutive steps (run k to group k mod ngroups)
    int run_shift;
};

// Adds the LPC per-face products of this lane's cell in slot order (the
// cell's ascending face order), taking only the slots in `use`.
template <bool FWD, int LPC, typename R>
__device__ __forceinline__ void add_cell_terms(const R (&p)[5], unsigned use, R (&D)[5], R (&acc)[5]) {
#define MFS_ADD_SLOT(S)                                                                      \
    if constexpr ((S) < LPC) {                                                               \
        R ps[5];                                                                             \
        _Pragma("unroll") for (int j = 0; j < 5; ++j) ps[j] = cell_bcast<LPC, (S)>(p[j]);    \
        if (use & (1u << (S))) {                                                             \
            _Pragma("unroll") for (int j = 0; j < 5; ++j) {                                  \
                if (FWD)                                                                     \
                    D[j] -= ps[j];                                                           \
                else                                                                         \
                    acc[j] += ps[j];                                                         \
            }                                                                                \
        }                                                                                    \
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

template <typename R>
__device__ __forceinline__ void load5r(const R* __restrict__ p, R v[5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) v[j] = p[j];
}

// This wave's staging area for a chunk's records (at most 80 words: 16
// cells of 5 values), in LDS.  Its lanes write their cells' values, then
// read back consecutive words; the wave barriers keep the compiler from
// moving the accesses across each other (a wave's LDS accesses complete in
// order).
__device__ __forceinline__ double* stage_words() {
    __shared__ double stage[kSweepBlock / 64][80];
    return stage[threadIdx.x / 64];
}

// The part of the face term that does not depend on the neighbour's DQ:
// the reference's FluxLUSGS3D split in two (FP64), or its increment form
// (FP32 operator of mixed precision).
template <typename R>
using FacePre = std::conditional_t<sizeof(R) == sizeof(double), FluxPre<R>, FluxDeltaPre<R>>;

// The chunk at position u of a sweep's order (the forward sweep walks the
// positions up, the backward sweep down).  Returns false when a wait
// expired (the error word is set).
//
// Latency: a chunk's operands come in three rounds of loads: its record;
// then the latest dependency's first value (lane 0, requested first), the
// cell's mask, slot neighbours and geometry, state, Diag and right-hand
// side; then the neighbours' states together with their DQ when the latest
// dependency was already published.  The DQ-independent half of every face
// term is evaluated before waiting, so only the DQ-dependent half, the
// sum, the limiter and the stores follow the dependency's arrival.
template <bool FWD, int LPC, typename R, bool THIN = true>
__device__ __forceinline__ bool sweep_chunk(const SweepArgs<R>& A, const FlowConst& K, int u) {
    using B = typename Real<R>::bits;
    constexpr B kSent = Real<R>::kSent;
    const int lane = (int)threadIdx.x & 63;
    const int k = lane / LPC;  // cell of the chunk
    const int s = lane % LPC;  // face slot
    R* const poll = FWD ? A.dqf : A.dqb;
    const int4 rec = A.crun[u];
    const int rep = FWD ? rec.z : rec.w;
    B rv = 0;
    if (lane == 0 && rep >= 0) rv = ld_dev<R>(poll + (size_t)rep * 5);
    const bool active = k < (rec.y & 0xff);
    // a thin level: every lane polls its own dependency (a batch interleaves
    // many solves, so no level of one solve is thin there: THIN = false)
    const bool thin = THIN && (rec.y & 0x100) != 0;
    const int c = rec.x + (active ? k : 0);  // idle lanes read the first cell and store nothing
    const unsigned m = (unsigned)A.depmask[c];
    unsigned kind = FWD ? (m & 0xffu) : ((m >> 8) & 0xffu);  // slots that take part
    const unsigned pendm = FWD ? kind : ((m >> 16) & 0xffu);  // slots with a DQ to wait for
    if (FWD && c == A.cell0) kind = 0;  // forward_lusgs.cpp:803-804: the first cell only divides
    const bool mine = active && ((kind >> s) & 1u);
    bool pend = mine && ((pendm >> s) & 1u);
    const size_t sl = (size_t)c * LPC + s;
    const int nb = A.snb[sl];
    const typename Real<R>::vec4 g = A.sgeo[sl];
    R qc[5], D[5];
    load5r(A.q + (size_t)c * 5, qc);
    const R dg = A.diag[c];
    if (FWD) {
#pragma unroll
        for (int j = 0; j < 5; ++j) D[j] = (R)A.res[(size_t)c * 5 + j];
    } else {
        load5r(A.dqf + (size_t)c * 5, D);
    }
    // the FP64 state the backward sweep updates (mixed precision), fetched
    // before the wait so that no memory round trip follows the dependencies
    double q64c[5];
    if constexpr (!FWD && sizeof(R) != sizeof(double)) load5(A.q64 + (size_t)c * 5, q64c);
    bool rep_seen = (rep < 0) || (__shfl(rv, 0, 64) != kSent);
    R qnb[5] = {R(0), R(0), R(0), R(0), R(0)};
    R dv[5] = {R(0), R(0), R(0), R(0), R(0)};
    if (mine) load5r(A.q + (size_t)nb * 5, qnb);
    const bool wide = wide_records(A.nq);  // the same for every wave
    // one read of every pending dependency (they are interior cells); a
    // value still the sentinel leaves the lane pending
    auto read_deps = [&] {
        R b[5];
        read_published(wide, static_cast<const R*>(poll), A.nC, nb, b);
        bool ok = true;
#pragma unroll
        for (int j = 0; j < 5; ++j) ok = ok && (__builtin_bit_cast(B, b[j]) != kSent);
        if (ok) {
#pragma unroll
            for (int j = 0; j < 5; ++j) dv[j] = b[j];
            pend = false;
        }
    };
    if (rep_seen && pend) read_deps();  // the dependencies are most likely out: read them now
    FacePre<R> pre;
    if (mine) {
        if constexpr (sizeof(R) == sizeof(double))
            flux_lusgs_pre<R>(pre, qnb, g.x, g.y, g.z, (R)K.gam, (R)K.p_bar, (R)K.lhs_omga);
        else
            flux_lusgs_delta_pre<R>(pre, qnb, g.x, g.y, g.z, (R)K.gam, (R)K.p_bar, (R)K.lhs_omga);
    }

    unsigned spins = 0;
    if (!rep_seen && !thin) {  // one lane waits for the chunk's latest dependency
        while (true) {
            B v = 0;
            if (lane == 0) v = ld_dev<R>(poll + (size_t)rep * 5);
            v = __shfl(v, 0, 64);
            if (v != kSent) break;
            if (++spins > kSpinLimit || wait_failed(A.err, spins)) {
                if (lane == 0) atomicExch(A.err, 1);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    // wide levels: the chunk's latest dependency is out, so the others most
    // likely are too; one read of each before the retry loop
    if (!thin && pend) read_deps();
    // every lane reads its face's dependency, retrying the rare unpublished one
    while (__ballot(pend) != 0) {
        if (pend) {
            B b[5];
            bool ok = true;
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                b[j] = ld_dev<R>(poll + (size_t)nb * 5 + j);
                ok = ok && (b[j] != kSent);
            }
            if (ok) {
#pragma unroll
                for (int j = 0; j < 5; ++j) dv[j] = __builtin_bit_cast(R, b[j]);
                pend = false;
            }
        }
        if (__ballot(pend) == 0) break;
        if (++spins > kSpinLimit || wait_failed(A.err, spins)) {
            if (lane == 0) atomicExch(A.err, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    // this lane's face term: forward 0.5*A*F (DQ -= term), backward A*F (summed)
    R p[5] = {R(0), R(0), R(0), R(0), R(0)};
    if (mine) {
        R fl[5];
        if constexpr (sizeof(R) == sizeof(double))  // the reference's expression, bit for bit
            flux_lusgs_post<R>(pre, dv, g.x, g.y, g.z, fl);
        else  // FP32 operator: the same function in increment form (no cancellation floor)
            flux_lusgs_delta_post<R>(pre, dv, g.x, g.y, g.z, fl);
        const R tmp = FWD ? R(0.5) * g.w : g.w;
#pragma unroll
        for (int j = 0; j < 5; ++j) p[j] = tmp * fl[j];
    }
    R acc[5] = {R(0), R(0), R(0), R(0), R(0)};
    add_cell_terms<FWD, LPC, R>(p, kind, D, acc);
    if (FWD) {
#pragma unroll
        for (int j = 0; j < 5; ++j) D[j] /= dg;
        if (c != A.cell0)
            dq_limit<R>(qc, D, (R)K.gamm1, (R)K.rho_min, (R)K.rho_max, (R)K.p_min, (R)K.p_max);
    } else {
        const R tmp = R(2.0) * dg;
#pragma unroll
        for (int j = 0; j < 5; ++j) D[j] -= acc[j] / tmp;
        dq_limit<R>(qc, D, (R)K.gamm1, (R)K.rho_min, (R)K.rho_max, (R)K.p_min, (R)K.p_max);
    }
    // Stores.  The chunk's cells are contiguous, so its 5-value records
    // form one run of words: lane w stores word w (and w + 64), and every
    // store instruction writes whole, consecutive words.  Stored value by
    // value from the cell's lanes, a record's words went out as 8- and
    // 32-byte pieces of lines: a quarter of the sweeps' write requests were
    // 32-byte partial writes, and the wide levels are bound by memory
    // traffic.  On thin levels the values are published straight from the
    // cell's lanes: there the staging round trip would lengthen the
    // level's dependent chain.
    const int nw = 5 * (rec.y & 0xff);  // words of the chunk's records
    {
        R* const own = FWD ? A.dqf : A.dqb;    // published
        R* const other = FWD ? A.dqb : A.dqf;  // re-armed for the next sweep
        if (thin) {
            if (active) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    if (j % LPC == s) st_dev<R>(own + (size_t)c * 5 + j, D[j]);
            }
        } else {
            double* const st = stage_words();
            if (active) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    if (j % LPC == s) st[k * 5 + j] = (double)D[j];  // exact for R = float too
            }
            __builtin_amdgcn_wave_barrier();
            for (int w = lane; w < nw; w += 64) st_dev<R>(own + (size_t)rec.x * 5 + w, (R)st[w]);
            __builtin_amdgcn_wave_barrier();
        }
        const R sent = __builtin_bit_cast(R, kSent);
        for (int w = lane; w < nw; w += 64) other[(size_t)rec.x * 5 + w] = sent;
    }
    if (!FWD) {
        double qnew[5], dd[5];
        if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll
            for (int j = 0; j < 5; ++j) qnew[j] = qc[j];
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) qnew[j] = q64c[j];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) dd[j] = (double)D[j];
        update_cell(qnew, dd, K);
        if (thin) {
            if (active) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    if (j % LPC == s) A.qout[(size_t)c * 5 + j] = qnew[j];
            }
        } else {
            double* const st = stage_words();
            if (active) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    if (j % LPC == s) st[k * 5 + j] = qnew[j];
            }
            __builtin_amdgcn_wave_barrier();
            for (int w = lane; w < nw; w += 64) A.qout[(size_t)rec.x * 5 + w] = st[w];
            __builtin_amdgcn_wave_barrier();
        }
    }
    return true;
}

template <bool FWD, int LPC, typename R>
__global__ __launch_bounds__(kSweepBlock) void k_sweep(SweepArgs<R> A, FlowConst K) {
    // a sweep that follows a failed one in the same stream does not start
    if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const int W = (int)blockDim.x / 64;
    const int w = (int)threadIdx.x / 64;
    if (A.flat) {
        // the other direction's counters are idle during this sweep
        if (blockIdx.x == 0 && threadIdx.x < kTicketGroups) A.tickets_next[threadIdx.x * kTicketStride] = 0;
        const int ng = A.ngroups;
        const int grp = (int)blockIdx.x % ng;
        int* const ctr = A.tickets + grp * kTicketStride;
        const int lane = (int)threadIdx.x & 63;
        int tk = 0;
        if (lane == 0) tk = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tk = __shfl(tk, 0, 64);
        while (true) {
            const int sh = A.run_shift;
            const int u = A.ubeg + ((((tk >> sh) * ng + grp) << sh) | (tk & ((1 << sh) - 1)));
            if (u >= A.uend) return;
            int next = 0;  // taken now, used after this chunk: its latency overlaps the chunk
            if (lane == 0) next = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!sweep_chunk<FWD, LPC, R>(A, K, FWD ? u : A.nchunks - 1 - u)) return;
            tk = __shfl(next, 0, 64);
        }
    }
    for (int kk = blockIdx.x; kk < A.ntiles; kk += gridDim.x) {
        const int k = FWD ? kk : A.ntiles - 1 - kk;
        const int c0 = A.tile_chunk[k], c1 = A.tile_chunk[k + 1];
        for (int j = w; j < c1 - c0; j += W) {
            const int t = FWD ? c0 + j : c1 - 1 - j;
            if (!sweep_chunk<FWD, LPC, R>(A, K, A.pos[t])) return;
        }
    }
}
