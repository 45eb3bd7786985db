// Device context of libmfs_hip.so (host-side definition).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/mfs.h"
#include "mfs_physics.hpp"

namespace mfs {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct KernelTimer {
    double ms = 0.0;
    int launches = 0;
};

}  // namespace mfs

struct mfs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;  // false for a solve of a batch (the batch's stream)
    int nC = 0, nF = 0, nB = 0, nN = 0;
    int precision = MFS_FP64;
    int maxf = 4;  // face slots per cell used by the kernels (4 or 6)
    int tile_cells = 4096, sweep_waves = 8;  // layout / launch options
    mfs_params P{};
    mfs::FlowConst K{};
    bool broken = false;

    std::vector<mfs::DevBuf> bufs;  // every allocation, freed at destroy
    size_t device_bytes = 0;

    // cell-major state (5 doubles per cell, ghosts at nC+k), two buffers
    double* q[2] = {nullptr, nullptr};
    int cur = 0;
    double *res = nullptr;  // 5 per cell
    double *dt = nullptr, *diag = nullptr;
    double *dqf = nullptr;  // forward-sweep DQ, 5 per cell
    double *dq = nullptr;   // final DQ, 5 per cell incl. ghosts (0)
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;  // reference layout: 5 blocks of nC+nB
    double* qn = nullptr;   // node values, 5 per device node
    // mixed precision (MFS_MIXED): FP32 implicit operator and gradients; the
    // FP64 arrays above keep the residual, time step and state
    float* q32 = nullptr;     // state snapshot of the iteration, 5 per cell incl. ghosts
    float* diag32 = nullptr;
    float* dqf32 = nullptr;   // forward-sweep DQ
    float* dq32 = nullptr;    // final DQ incl. ghosts (0)
    float *gx32 = nullptr, *gy32 = nullptr, *gz32 = nullptr;
    float* qn32 = nullptr;
    float* ell_w32 = nullptr; // node weights
    float4* fgeo32 = nullptr; // unit normal + area
    // faces (device order)
    int2* fcell = nullptr;
    double4* fgeo = nullptr;
    double2* fdn = nullptr;
    int *fno = nullptr, *fn = nullptr, *btype = nullptr;
    // cells
    int *cfo = nullptr, *cf = nullptr;
    double* vol = nullptr;
    // node gather
    int* nlo = nullptr;
    int* ell_off = nullptr;   // per wave of 64 nodes: first entry (entry-major rows of 64)
    int2* ell_ab = nullptr;
    double* ell_w = nullptr;
    double* nden = nullptr;
    int nsym = 0;
    int* sym_node = nullptr;
    double* sym_n = nullptr;
    std::vector<int> node_of_dev;  // host: device node -> mesh node
    std::vector<int> cell_of_dev;  // host: device cell -> mesh cell
    // sweeps
    int4* crun = nullptr;       // per position of the flat order: first cell, count, rep fwd, rep bwd
    int* pos = nullptr;         // per chunk (device order): its position in the flat order
    int* tile_chunk = nullptr;  // ntiles+1: first chunk of each tile
    int* snb = nullptr;         // lpc per device cell: neighbour across the slot's face
    double4* sgeo = nullptr;    // lpc per device cell: normal away from the cell, area
    float4* sgeo32 = nullptr;   // the same in FP32 (mixed precision)
    int sweep_by_tiles = 0;
    int ntiles = 0;
    int* depmask = nullptr;
    int* tickets = nullptr;  // ticket counters of the flat sweep schedule: forward, then backward
    int* gface = nullptr;    // nB: device face of each ghost cell nC+k
    int* snode = nullptr;    // tetrahedral meshes: 3 face nodes per slot, 12 per cell
    double* sdn = nullptr;   // per slot: |(x_fc - x_cc).n| on this cell's side
    int* sflip = nullptr;    // per cell: slots whose face has this cell second (bits 0-7), face count (8-11)
    int ghost_refresh = 0;   // SetGhostVariables after each update in mfs_step
    // mfs_step's gradient on a second stream (gradient_stream options 2-5):
    // iteration k's node values and gradient run beside iteration k+1's
    // residual pass and sweeps.  sdone: recorded on `stream` where the
    // gradient starts; gdone[k & 1]: on `gstream` after iteration k's gradient, which
    // reads the state buffer that the backward sweep of iteration k+2
    // overwrites; gpend[j]: gdone[j] recorded and not yet waited on.
    hipStream_t gstream = nullptr;
    hipEvent_t sdone = nullptr, gdone[2] = {nullptr, nullptr};
    bool gpend[2] = {false, false};
    int gradient_stream = 1;  // 1: on the solver stream (fused with the next residual pass), 2-5: second stream
    int cell0 = 0;  // device id of mesh cell 0
    int nchunk = 0, lpc = 4;
    int nlev = 0;
    int sweep_grid = 0;  // resident workgroups of the sweep launch
    // The sweeps' thin tail: positions [tail_u, nchunk) of the flat order
    // are the longest run of thin levels (at most kThinLevelChunks chunks
    // each) that ends the order.  With tail_grid > 0 each sweep runs as two
    // launches, its wide levels on sweep_grid workgroups and its tail on
    // tail_grid, so that other work can fill the CUs the tail leaves idle.
    int tail_u = 0;
    int tail_grid = 0;
    int tail_cells = 0;  // cells of the chunks at positions [tail_u, nchunk)
    int compute_units = 0;  // CUs of the device
    // reductions and history
    double* part = nullptr;  // 8 * nblk, value-major: 5 squared-norm partials, dt min, dt max, spare
    double* run = nullptr;   // [0] dt_max [1] dt_min [2] clamp flag [3..7] last norms
    int* err = nullptr;      // [0] sweep wait limit hit
    double* hist = nullptr;
    int hist_cap = 0, hist_n = 0;
    int nblk_cell = 0;

    // host mirrors
    int iter_done = 0;
    double cfl_min = 0.0;

    // timing
    int timing = 0;  // 0 off, 1 every kernel group, 2 the sweeps only
    std::map<std::string, mfs::KernelTimer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;  // timed intervals
    std::vector<hipEvent_t> pending_events;  // the events they use, each once (intervals may share one)
    std::vector<hipEvent_t> event_pool;
};
