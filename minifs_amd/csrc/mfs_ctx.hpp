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
    int nC = 0, nF = 0, nB = 0, nN = 0;
    int precision = MFS_FP64;
    mfs_params P{};
    mfs::FlowConst K{};
    bool broken = false;

    std::vector<mfs::DevBuf> bufs;  // every allocation, freed at destroy
    size_t device_bytes = 0;

    // state (two buffers, 5 blocks of nC+nB each) and outputs
    double* q[2] = {nullptr, nullptr};
    int cur = 0;
    double *res = nullptr, *dt = nullptr, *diag = nullptr, *dq = nullptr;
    double *gx = nullptr, *gy = nullptr, *gz = nullptr, *qn = nullptr;
    // faces
    int2* fcell = nullptr;
    double4* fgeo = nullptr;
    double2* fdn = nullptr;
    int *fno = nullptr, *fn = nullptr, *btype = nullptr;
    // cells
    int *cfo = nullptr, *cf = nullptr;
    double* vol = nullptr;
    // node gather
    int* nlo = nullptr;
    int2* nlab = nullptr;
    double *nlw = nullptr, *nden = nullptr;
    int nsym = 0;
    int* sym_node = nullptr;
    double* sym_n = nullptr;
    // sweeps
    int *sched_f = nullptr, *sched_b = nullptr, *flags = nullptr;
    unsigned int* tickets = nullptr;  // [0] forward, [1] backward
    int nlev_f = 0, nlev_b = 0;
    int epoch = 0;
    // reductions and history
    double* part = nullptr;  // nblk * 8: 5 squared-norm partials, dt min, dt max, spare
    double* run = nullptr;   // [0] dt_max [1] dt_min [2] clamp flag [3..7] last norms
    int* err = nullptr;      // [0] sweep spin limit hit
    double* hist = nullptr;
    int hist_cap = 0, hist_n = 0;
    int nblk_cell = 0;

    // host mirrors
    int iter_done = 0;
    double cfl_min = 0.0;

    // timing
    bool timing = false;
    std::map<std::string, mfs::KernelTimer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
};
