// HIP kernels and device context of the MI355X solver path (gfx950).
//
// Every reference loop that scatters face contributions into cells is run
// here as a per-cell gather over the cell's faces in ascending face id,
// which adds the same terms in the same order as the reference's serial
// face loop: results are bitwise equal, with no atomics and no colouring.
// The device layout (cell-major state, first-touch face and node storage,
// level schedules for the sweeps) is built on the host by mfs_prep.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <mutex>
#include <sstream>
#include <vector>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>

#include "mfs_ctx.hpp"
#include "mfs_mesh.hpp"
#include "mfs_grid.hpp"
#include "mfs_prep.hpp"

#include <hipcub/hipcub.hpp>

namespace mfs {
void set_error(const std::string& msg);  // mfs_capi.cpp
}

// The static scratch of mfs_step_noalloc: per-block partials of the
// residual norms and time-step extrema for contexts of up to
// kNoAllocMaxCells cells, then the two sweeps' ticket counters and the
// side-stream gradient's block counters (2.03 MiB).
constexpr int kNoAllocMaxCells = 1 << 23;
constexpr size_t kNoAllocPartDoubles = (size_t)(((kNoAllocMaxCells + 255) / 256 + 7) / 8 * 8) * 8;
constexpr size_t kNoAllocTicketInts = 2 * 64 * 64 + 2 * 8 * 64;
constexpr size_t kNoAllocBytes = kNoAllocPartDoubles * sizeof(double) + kNoAllocTicketInts * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace mfs {
namespace dev {

constexpr int kBlock = 256;
constexpr int kXcd = 8;  // XCDs of an MI355X; cell-kernel grids are a multiple of this
constexpr int kSweepWave = 64;
constexpr int kSweepBlock = 1024;  // largest sweep workgroup (16 waves); one tile per workgroup
// Resident sweep waves per CU.  The sweeps need every wave of the grid
// resident at once, and the occupancy query is not a safe bound: for the
// FP32 sweeps (72 VGPRs, 103-107 SGPRs) it reports 7 workgroups of 4 waves
// per CU while an MI355X holds only 6 of them at once (the SGPRs are the
// limit), and such a grid waits on chunks owned by waves that never start
// until its bounded waits expire.  16 waves per CU (4 per SIMD) were all
// resident for both sweep footprints, and more sweep waves do not run
// faster.
//
// The sweeps are plain launches of that resident-sized grid.  They never
// synchronise the whole grid, so a cooperative launch would add only its
// size check, and it takes a device-wide wave-sync resource that one
// process at a time can hold (a second process's cooperative launch
// fails).  The context's stream runs the solver's kernels one after
// another, so a sweep starts on an otherwise idle device unless another
// process shares the GPU; if that keeps part of the grid from starting,
// the bounded waits end the sweep with an error instead of a hang.
constexpr int kSweepWavesPerCU = 16;
constexpr unsigned kSpinLimit = 1u << 23;  // bound on the waits of one sweep wave
// Ticket groups of the flat sweep schedule.  With n groups in use (64, or
// fewer when the grid has fewer workgroups), position u of the flat order
// belongs to group u mod n, and the waves of the workgroups b with
// b mod n = g take group g's positions in order from the group's counter.  One counter for the whole grid would serialise every chunk on
// one address (about 90 M atomics/s measured on an MI355X, 4.7 ms per
// sweep of 6M cells); 64 counters cost well under a tenth of a sweep.
constexpr int kTicketGroups = 64;
constexpr int kTicketStride = 64;  // ints between counters (256 bytes)
// The counters of a context or a batch: the two sweeps' ticket sets, then
// one block counter per XCD for each of two claiming kernels on the second
// stream (a context's gradient; a batch's node values and gradient).
constexpr size_t kGradClaimOffset = 2 * (size_t)kTicketGroups * kTicketStride;
constexpr size_t kNodeClaimOffset = kGradClaimOffset + (size_t)dev::kXcd * kTicketStride;
constexpr size_t kTicketInts = kNodeClaimOffset + (size_t)dev::kXcd * kTicketStride;
// Wide launches hand out runs of 4 consecutive steps per ticket group: on
// the 6.0M-cell mesh 1.2% faster per FP64 iteration than single steps
// (1.0% mixed), and 8-step runs gained less.  The thin tail keeps single
// steps: its levels are a few chunks each.
constexpr int kWideRunShift = 2;
constexpr unsigned kClampBlocks = 1024;  // grid of the ratio_dtmax clamp (grid-stride over the cells)
// Where mfs_step runs CalculateGradient: 1 on the solver stream, fused with
// the next iteration's residual pass; 2, 3, 4, 5 on a second stream,
// started after the update, after the next residual pass, after the next
// forward sweep, after the next forward sweep's wide levels.
// Measured on 6.0M cells, ms per iteration, with one launch per sweep
// (modes 1/2/3/4): FP64 4.47/4.51/4.23/4.50, mixed precision
// 3.63/3.66/3.55/3.34; with the thin tail launched apart (modes 1/3/4/5):
// FP64 4.44/4.14/4.15/4.04, mixed 3.47 (3), 3.33 (4), 3.33 (5).  Mode 5
// runs the node values and the gradient in the CUs that the two sweeps'
// thin tails leave idle.
constexpr int kDefaultGradientStreamFP64 = 5;
constexpr int kDefaultGradientStreamMixed = 5;
// The same for the batch step (1: fused with the next residual pass; 2, 4:
// second stream after the update / after the next forward sweep).  64
// seeded solves (21.9M cells), ms per batch iteration, modes 1/2/4: FP64
// 14.20/14.59/14.52, mixed precision 11.73/12.00/11.47.
constexpr int kDefaultBatchGradientStreamFP64 = 1;
constexpr int kDefaultBatchGradientStreamMixed = 4;
// The sweeps' thin tail runs as a launch of its own on this many
// workgroups (a thin level has at most 64 chunks), when it holds at least
// kMinTailChunks chunks.  6.0M cells, FP64, gradient mode 5, ms per
// iteration for 32/64/128/256 workgroups: 4.08/4.04/4.05/4.05.
constexpr int kDefaultTailGrid = 64;
constexpr int kMinTailChunks = 256;

#include "mfs_kernels.hpp"
#include "mfs_sweep.hpp"
#include "mfs_batch.hpp"
#include "mfs_setup.hpp"

}  // namespace dev

// ===========================================================================
// ComputeMetrics on the current device (the kernels of mfs_setup.hpp); the
// results land in the grid's host arrays, bitwise those of compute_metrics.
namespace {

struct DevScratch {
    std::vector<void*> bufs;
    hipStream_t stream = nullptr;
    ~DevScratch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* b : bufs) (void)hipFree(b);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return static_cast<T*>(p);
    }
};

}  // namespace

std::string compute_metrics_device(Grid& g) {
    const RawMesh& m = g.m;
    const int nC = m.nC, nB = m.nB, nF = m.nF, nN = m.nN, n = nC + nB;
    if (nC <= 0 || nF <= 0) return "empty mesh";
    const std::vector<int> fno = m.f2n_offsets();
    DevScratch d;
    if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess) return "HIP stream creation failed";
    hipStream_t st = d.stream;
    dev::MetricArgs A{};
    A.nC = nC;
    A.nB = nB;
    A.nF = nF;
    int* dfno = d.alloc<int>(fno.size());
    int* df2n = d.alloc<int>(m.f2n.size());
    int* df2c = d.alloc<int>(m.f2c.size());
    double* dx = d.alloc<double>(nN);
    double* dy = d.alloc<double>(nN);
    double* dz = d.alloc<double>(nN);
    int* cnt = d.alloc<int>((size_t)nC + 1);
    int* cfo = d.alloc<int>((size_t)nC + 1);
    int* cursor = d.alloc<int>(nC);
    int* cf = d.alloc<int>(2 * (size_t)nF - nB);
    double* fbuf = d.alloc<double>(7 * (size_t)nF);
    double* cbuf = d.alloc<double>(4 * (size_t)n + 3 * (size_t)nC);
    int* small = d.alloc<int>(1);
    if (!dfno || !df2n || !df2c || !dx || !dy || !dz || !cnt || !cfo || !cursor || !cf || !fbuf || !cbuf || !small)
        return "device memory allocation failed";
    bool ok = hipMemcpyAsync(dfno, fno.data(), fno.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(df2n, m.f2n.data(), m.f2n.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(df2c, m.f2c.data(), m.f2c.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(dx, m.x.data(), nN * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(dy, m.y.data(), nN * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(dz, m.z.data(), nN * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemsetAsync(cnt, 0, ((size_t)nC + 1) * sizeof(int), st) == hipSuccess &&
              hipMemsetAsync(cursor, 0, (size_t)nC * sizeof(int), st) == hipSuccess &&
              hipMemsetAsync(small, 0, sizeof(int), st) == hipSuccess;
    if (!ok) return "upload failed";
    const unsigned gF = (unsigned)((nF + dev::kBlock - 1) / dev::kBlock);
    const unsigned gC = (unsigned)((nC + dev::kBlock - 1) / dev::kBlock);
    const unsigned gB = (unsigned)std::max(1, (nB + dev::kBlock - 1) / dev::kBlock);
    // cell -> faces in ascending face id
    hipLaunchKernelGGL(dev::k_c2f_count, dim3(gF), dim3(dev::kBlock), 0, st, nF, nB, df2c, cnt);
    size_t tmp_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, cfo, nC + 1, st) != hipSuccess) return "scan failed";
    void* tmp = d.alloc<char>(tmp_bytes);
    if (!tmp || hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, cfo, nC + 1, st) != hipSuccess)
        return "scan failed";
    hipLaunchKernelGGL(dev::k_c2f_fill, dim3(gF), dim3(dev::kBlock), 0, st, nF, nB, df2c, cfo, cursor, cf);
    hipLaunchKernelGGL(dev::k_c2f_sort, dim3(gC), dim3(dev::kBlock), 0, st, nC, cfo, cf);
    A.fno = dfno;
    A.f2n = df2n;
    A.f2c = df2c;
    A.x = dx;
    A.y = dy;
    A.z = dz;
    A.cfo = cfo;
    A.cf = cf;
    A.xfc = fbuf;
    A.yfc = fbuf + (size_t)nF;
    A.zfc = fbuf + 2 * (size_t)nF;
    A.xfn = fbuf + 3 * (size_t)nF;
    A.yfn = fbuf + 4 * (size_t)nF;
    A.zfn = fbuf + 5 * (size_t)nF;
    A.area = fbuf + 6 * (size_t)nF;
    A.xcc = cbuf;
    A.ycc = cbuf + (size_t)n;
    A.zcc = cbuf + 2 * (size_t)n;
    A.vol = cbuf + 3 * (size_t)n;
    A.xcct = cbuf + 4 * (size_t)n;
    A.ycct = A.xcct + nC;
    A.zcct = A.ycct + nC;
    A.small = small;
    hipLaunchKernelGGL(dev::k_face_avg, dim3(gF), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_cell_avg, dim3(gC), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_face_geom, dim3(gF), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_cell_vol, dim3(gC), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_cell_mark, dim3(gC), dim3(dev::kBlock), 0, st, A);
    if (nB > 0) hipLaunchKernelGGL(dev::k_ghost_geom, dim3(gB), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_face_normal_fix, dim3(gF), dim3(dev::kBlock), 0, st, A);
    hipLaunchKernelGGL(dev::k_cell_centroid_fix, dim3(gC), dim3(dev::kBlock), 0, st, A, centroid_scale());
    if (hipGetLastError() != hipSuccess) return "metric kernel launch failed";
    int nsmall = 0;
    g.xfc.resize(nF);
    g.yfc.resize(nF);
    g.zfc.resize(nF);
    g.xfn.resize(nF);
    g.yfn.resize(nF);
    g.zfn.resize(nF);
    g.area.resize(nF);
    g.xcc.resize(n);
    g.ycc.resize(n);
    g.zcc.resize(n);
    g.vol.resize(n);
    std::vector<double>* fo[7] = {&g.xfc, &g.yfc, &g.zfc, &g.xfn, &g.yfn, &g.zfn, &g.area};
    std::vector<double>* co[4] = {&g.xcc, &g.ycc, &g.zcc, &g.vol};
    ok = hipMemcpyAsync(&nsmall, small, sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess;
    for (int k = 0; k < 7 && ok; ++k)
        ok = hipMemcpyAsync(fo[k]->data(), fbuf + k * (size_t)nF, nF * sizeof(double), hipMemcpyDeviceToHost, st) ==
             hipSuccess;
    for (int k = 0; k < 4 && ok; ++k)
        ok = hipMemcpyAsync(co[k]->data(), cbuf + k * (size_t)n, n * sizeof(double), hipMemcpyDeviceToHost, st) ==
             hipSuccess;
    if (!ok || hipStreamSynchronize(st) != hipSuccess) return "metric download failed";
    if (nsmall > 0) {
        std::ostringstream os;
        os << "There are " << nsmall << " cell's vol too small! This must correct!";
        return os.str();
    }
    return "";
}

// ===========================================================================
// Context management and launch sequences.
namespace {

using dev::kBlock;
using dev::kSweepWave;
using dev::kSweepWavesPerCU;
using dev::kTicketGroups;
using dev::kTicketStride;
using dev::kGradClaimOffset;
using dev::kNodeClaimOffset;
using dev::kTicketInts;

bool hip_ok(mfs_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    if (ctx) ctx->broken = true;
    return false;
}

template <typename T>
bool dalloc(mfs_ctx* ctx, T*& p, size_t count, const char* what) {
    void* v = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (!hip_ok(ctx, hipMalloc(&v, bytes), what)) return false;
    ctx->bufs.push_back({v, bytes});
    ctx->device_bytes += bytes;
    p = static_cast<T*>(v);
    return true;
}

template <typename T>
bool upload(mfs_ctx* ctx, T* dst, const T* src, size_t count, const char* what) {
    if (count == 0) return true;
    return hip_ok(ctx, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream), what);
}

inline unsigned nblk(size_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }
// grid of a kernel that uses xcd_block(): whole multiples of kXcd blocks
inline unsigned nblk_xcd(size_t n) { return (nblk(n) + dev::kXcd - 1) / dev::kXcd * dev::kXcd; }

// Event bracketing of one launch group when timing is enabled.
// Events recorded on the stream before and after the launches of a scope
// (timing level 1; the sweeps time themselves, SweepTimes below).
struct Timed {
    mfs_ctx* ctx;
    const char* name;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(mfs_ctx* c, const char* nm, hipStream_t s = nullptr) : ctx(c), name(nm), st(s ? s : c->stream) {
        if (ctx->timing != 1) return;
        a = take();
        b = take();
        if (!a || !b) a = b = nullptr;
        if (a) (void)hipEventRecord(a, st);
    }
    ~Timed() {
        if (ctx->timing != 1 || !a || !b) return;
        (void)hipEventRecord(b, st);
        ctx->pending.push_back({name, {a, b}});
        ctx->pending_events.push_back(a);
        ctx->pending_events.push_back(b);
    }
    hipEvent_t take() {
        if (!ctx->event_pool.empty()) {
            hipEvent_t e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};

void collect_timing(mfs_ctx* ctx) {
    for (auto& pe : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            auto& t = ctx->timers[pe.first];
            t.ms += ms;
            t.launches += 1;
        }
    }
    ctx->pending.clear();
    ctx->event_pool.insert(ctx->event_pool.end(), ctx->pending_events.begin(), ctx->pending_events.end());
    ctx->pending_events.clear();
}

// The timing of one sweep: events attached to its launches (hipExtLaunchKernel),
// e[0], e[1] around its first launch and, when it is split, e[2], e[3]
// around its second.  It times the sweep from its first launch's start to
// its last launch's end and, split, each launch on its own.
struct SweepTimes {
    mfs_ctx* ctx;
    int n = 0;
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    SweepTimes(mfs_ctx* c, bool split) : ctx(c) {
        if (!ctx->timing) return;
        const int want = split ? 4 : 2;
        for (int i = 0; i < want; ++i) {
            hipEvent_t ev = nullptr;
            if (!ctx->event_pool.empty()) {
                ev = ctx->event_pool.back();
                ctx->event_pool.pop_back();
            } else if (hipEventCreate(&ev) != hipSuccess) {
                ev = nullptr;
            }
            if (!ev) break;
            e[n++] = ev;
        }
        if (n != want) {  // no timing rather than a partial one
            for (int i = 0; i < n; ++i) ctx->event_pool.push_back(e[i]);
            n = 0;
            e[0] = e[1] = e[2] = e[3] = nullptr;
        }
    }
    hipEvent_t start(int launch) const { return n ? e[2 * launch] : nullptr; }
    hipEvent_t stop(int launch) const { return n ? e[2 * launch + 1] : nullptr; }
    // after the launches: the names of the whole sweep and of its two launches
    void done(const char* sweep, const char* first, const char* second) {
        if (!n) return;
        ctx->pending.push_back({sweep, {e[0], e[n - 1]}});
        if (n == 4) {
            ctx->pending.push_back({first, {e[0], e[1]}});
            ctx->pending.push_back({second, {e[2], e[3]}});
        }
        for (int i = 0; i < n; ++i) ctx->pending_events.push_back(e[i]);
        n = 0;
    }
    ~SweepTimes() {  // an error path left the events unused
        for (int i = 0; i < n; ++i) ctx->event_pool.push_back(e[i]);
    }
};

bool check_launch(mfs_ctx* ctx, const char* what) { return hip_ok(ctx, hipGetLastError(), what); }

// Sweeps of different streams must not be in flight together on one device.
// A sweep grid is sized to be resident on an otherwise idle device; two
// grids that each hold part of the GPU could each wait on chunks owned by
// waves the other keeps from starting, until their bounded waits end both
// sweeps with an error.  So every sweep pair (forward and backward, of a
// context or of a batch, from any host thread) is ordered after the
// previous one launched on the same device: under a per-device lock, a
// stream that is not the one that launched the device's last sweep pair
// records an event on that stream (which then follows its sweeps, enqueued
// under the same lock) and waits on it on the GPU.  A stream that sweeps
// after itself needs no event: each event record costs the stream about
// 7 us of idle GPU between two kernels.  Kernels that do not wait on other
// waves (everything but the sweeps) may still overlap.
// The table is allocated once and never freed: static destructors run at
// process exit, after which the HIP runtime may be gone.
constexpr int kMaxDevices = 64;
struct SweepOrder {
    std::mutex mu;
    hipEvent_t last = nullptr;          // recorded on last_stream when another stream takes its turn
    hipStream_t last_stream = nullptr;  // stream of the device's latest sweep pair (null: none yet)
};
// last_stream of a table whose latest sweeping stream was destroyed: `last`
// was recorded on that stream before, and follows its sweeps.
const hipStream_t kGoneStream = reinterpret_cast<hipStream_t>(~uintptr_t(0));
SweepOrder* const g_sweep_order = new SweepOrder[kMaxDevices];

// Creates the device's ordering event, once; called with `device` current
// when a context or a batch is created there, so that stepping never
// creates one.
bool sweep_order_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device id out of range for the sweep order table");
        return false;
    }
    SweepOrder& o = g_sweep_order[device];
    std::lock_guard<std::mutex> lk(o.mu);
    if (o.last) return true;
    if (hipEventCreateWithFlags(&o.last, hipEventDisableTiming) != hipSuccess) {
        o.last = nullptr;
        set_error("hipEventCreate(sweep order)");
        return false;
    }
    return true;
}

// Holds the device's lock from before the sweeps are enqueued until they
// are.
class SweepTurn {
  public:
    SweepTurn(int device, hipStream_t stream) : o_(g_sweep_order[device & (kMaxDevices - 1)]), st_(stream), lk_(o_.mu) {
        if (device < 0 || device >= kMaxDevices) {
            err_ = "device id out of range for the sweep order table";
            return;
        }
        if (!o_.last) {
            err_ = "sweep order event of the device not created";
            return;
        }
        if (!o_.last_stream || o_.last_stream == st_) return;
        if (o_.last_stream != kGoneStream && hipEventRecord(o_.last, o_.last_stream) != hipSuccess)
            err_ = "hipEventRecord(sweep order)";
        else if (hipStreamWaitEvent(st_, o_.last, 0) != hipSuccess)
            err_ = "hipStreamWaitEvent(sweep order)";
    }
    // after the sweeps are enqueued
    bool done() {
        if (err_) return false;
        o_.last_stream = st_;
        return true;
    }
    const char* error() const { return err_; }

  private:
    SweepOrder& o_;
    hipStream_t st_;
    std::lock_guard<std::mutex> lk_;
    const char* err_ = nullptr;
};

bool sweep_turn_failed(mfs_ctx* ctx, const char* what) {
    set_error(what);
    if (ctx) ctx->broken = true;
    return false;
}

// A stream about to be destroyed leaves the table, so that a later stream
// created at the same address still waits for the device's last sweeps.
void forget_sweep_stream(int device, hipStream_t stream) {
    if (device < 0 || device >= kMaxDevices) return;
    SweepOrder& o = g_sweep_order[device];
    std::lock_guard<std::mutex> lk(o.mu);
    if (o.last_stream != stream) return;
    (void)hipEventRecord(o.last, stream);  // follows the stream's sweeps
    o.last_stream = kGoneStream;
}

// ---------------------------------------------------------------------------
// mfs_step_noalloc's scratch: the per-block partials of the residual norms
// and time-step extrema (8 doubles per cell block) and the two sets of
// sweep ticket counters, in the static device array `scratch` declared at
// file scope below, sized for contexts of up to kNoAllocMaxCells cells
// (2.03 MiB; every GPU the library runs on holds its own copy until the
// process ends).  In host code `scratch` names a host-side stand-in: the
// device address of each GPU's copy is taken once per device, with that
// device current, by hipGetSymbolAddress at context creation and kept in
// g_noalloc_scratch; only that address is passed to kernels.
static_assert(kNoAllocPartDoubles ==
                  (size_t)(((kNoAllocMaxCells + dev::kBlock - 1) / dev::kBlock + dev::kXcd - 1) / dev::kXcd * dev::kXcd) * 8,
              "scratch partials: 8 doubles per cell block of the largest context");
static_assert(kNoAllocTicketInts == kTicketInts, "scratch tickets: both sweeps and the gradient's counters");
static_assert(7 * (kNoAllocPartDoubles / 8) + 7 * (size_t)dev::kReduceBlock <= kNoAllocPartDoubles,
              "scratch partials: room for the reduction's folds after the largest context's records");

std::mutex g_noalloc_mu;
void* g_noalloc_scratch[kMaxDevices] = {};

// With `device` current: takes the address of this device's copy of
// `scratch`, once per device.
bool noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device id out of range for the scratch table");
        return false;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_scratch[device]) return true;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return false;
    }
    g_noalloc_scratch[device] = ptr;
    return true;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_scratch[device];
}

double cfl_of_iter(const mfs_params& P, int iter_done) {
    // LimitTimeStep (compute_timestep.cpp:868-871)
    if (iter_done > P.cfl_nstep) return P.cfl_end;
    const double ratio = P.cfl_end / P.cfl_start;
    return P.cfl_start * pow(ratio, (double)iter_done / P.cfl_nstep);
}

bool ensure_hist(mfs_ctx* ctx, int need) {
    if (need <= ctx->hist_cap) return true;
    int cap = std::max(need, std::max(64, 2 * ctx->hist_cap));
    void* v = nullptr;
    if (!hip_ok(ctx, hipMalloc(&v, sizeof(double) * 7 * (size_t)cap), "hipMalloc(history)")) return false;
    double* nh = static_cast<double*>(v);
    if (ctx->hist && ctx->hist_n > 0 &&
        !hip_ok(ctx,
                hipMemcpyAsync(nh, ctx->hist, sizeof(double) * 7 * ctx->hist_n, hipMemcpyDeviceToDevice, ctx->stream),
                "copy history"))
        return false;
    if (ctx->hist) {
        if (!hip_ok(ctx, hipStreamSynchronize(ctx->stream), "history sync")) return false;
        (void)hipFree(ctx->hist);
    }
    ctx->hist = nh;
    ctx->hist_cap = cap;
    return true;
}

// ---- launch helpers
template <int MAXF>
bool launch_flux_t(mfs_ctx* ctx, bool accum) {
    Timed t(ctx, "flux");
    const double* q = ctx->q[ctx->cur];
    if (accum)
        hipLaunchKernelGGL((dev::k_flux<true, MAXF>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, q, ctx->res, ctx->part, ctx->K);
    else
        hipLaunchKernelGGL((dev::k_flux<false, MAXF>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, q, ctx->res, ctx->part, ctx->K);
    return check_launch(ctx, "k_flux");
}
bool launch_flux(mfs_ctx* ctx, bool accum) {
    return ctx->maxf == 4 ? launch_flux_t<4>(ctx, accum) : launch_flux_t<6>(ctx, accum);
}

template <int MAXF>
bool launch_timestep_t(mfs_ctx* ctx, bool with_diag, double cfl, double cfl_min) {
    Timed t(ctx, "timestep");
    const double* q = ctx->q[ctx->cur];
    if (with_diag)
        hipLaunchKernelGGL((dev::k_timestep<true, MAXF>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, ctx->fdn, q, ctx->vol, ctx->dt, ctx->diag, ctx->part,
                           ctx->K, cfl, cfl_min);
    else
        hipLaunchKernelGGL((dev::k_timestep<false, MAXF>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, ctx->fdn, q, ctx->vol, ctx->dt, ctx->diag, ctx->part,
                           ctx->K, cfl, cfl_min);
    return check_launch(ctx, "k_timestep");
}
bool launch_timestep(mfs_ctx* ctx, bool with_diag, double cfl, double cfl_min) {
    return ctx->maxf == 4 ? launch_timestep_t<4>(ctx, with_diag, cfl, cfl_min)
                          : launch_timestep_t<6>(ctx, with_diag, cfl, cfl_min);
}

template <int MAXF>
bool launch_flux_ts_t(mfs_ctx* ctx, double cfl, double cfl_min) {
    Timed t(ctx, "flux_ts");
    constexpr int LPC = MAXF == 4 ? 4 : 8;  // slots per cell of the slot tables
    if (ctx->precision == MFS_MIXED)
        hipLaunchKernelGGL((dev::k_flux_ts_slot<LPC, true>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream,
                           ctx->nC, ctx->snb, ctx->sgeo, ctx->sdn, ctx->sflip, ctx->q[ctx->cur], ctx->vol, ctx->res,
                           ctx->dt, ctx->diag, ctx->part, ctx->K, cfl, cfl_min, ctx->q32, ctx->diag32);
    else
        hipLaunchKernelGGL((dev::k_flux_ts_slot<LPC, false>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream,
                           ctx->nC, ctx->snb, ctx->sgeo, ctx->sdn, ctx->sflip, ctx->q[ctx->cur], ctx->vol, ctx->res,
                           ctx->dt, ctx->diag, ctx->part, ctx->K, cfl, cfl_min, nullptr, nullptr);
    return check_launch(ctx, "k_flux_ts_slot");
}
bool launch_flux_ts(mfs_ctx* ctx, double cfl, double cfl_min) {
    return ctx->maxf == 4 ? launch_flux_ts_t<4>(ctx, cfl, cfl_min) : launch_flux_ts_t<6>(ctx, cfl, cfl_min);
}

// Mixed precision, per-entry-point path: the FP32 operator inputs from the
// current state and Diag (mfs_step writes them in the fused kernel).
bool launch_mixed_inputs(mfs_ctx* ctx) {
    hipLaunchKernelGGL(dev::k_mixed_inputs, dim3(nblk(ctx->nC)), dim3(kBlock), 0, ctx->stream, ctx->nC,
                       ctx->q[ctx->cur], ctx->diag, ctx->q32, ctx->diag32);
    return check_launch(ctx, "k_mixed_inputs");
}

// Above this many cell blocks the reduction runs as k_reduce_fold +
// k_reduce_tree (the same bits as k_reduce, from many CUs).
constexpr int kFoldMinBlocks = 2048;
// The folds of k_reduce_fold go to the partials' buffer from its spare
// row (index 7) on: 7 * kReduceBlock doubles.
size_t partial_doubles(int nblk) {
    return std::max<size_t>(8 * (size_t)nblk, 7 * (size_t)nblk + 7 * (size_t)dev::kReduceBlock);
}

bool launch_reduce(mfs_ctx* ctx, int mode, int slot) {
    Timed t(ctx, "reduce");
    const int nb = (int)ctx->nblk_cell;
    if (nb > kFoldMinBlocks) {
        double* const fold = ctx->part + (size_t)7 * nb;
        hipLaunchKernelGGL(dev::k_reduce_fold, dim3(dev::kReduceBlock * dev::kFoldLanes / dev::kFoldThreads),
                           dim3(dev::kFoldThreads), 0, ctx->stream, nb, ctx->part, fold);
        if (!check_launch(ctx, "k_reduce_fold")) return false;
        hipLaunchKernelGGL(dev::k_reduce_tree, dim3(1), dim3(dev::kReduceBlock), 0, ctx->stream, fold, ctx->run,
                           ctx->hist, slot, mode, ctx->P.ratio_dtmax);
        return check_launch(ctx, "k_reduce_tree");
    }
    hipLaunchKernelGGL(dev::k_reduce, dim3(1), dim3(dev::kReduceBlock), 0, ctx->stream, nb, ctx->part, ctx->run,
                       ctx->hist, slot, mode, ctx->P.ratio_dtmax);
    return check_launch(ctx, "k_reduce");
}

template <int MAXF>
bool launch_diag_t(mfs_ctx* ctx, bool clamp_only) {
    Timed t(ctx, clamp_only ? "clamp" : "diag");
    if (clamp_only)  // usually off: a small grid that returns at once
        hipLaunchKernelGGL((dev::k_clamp<MAXF>), dim3(std::min<unsigned>(ctx->nblk_cell, dev::kClampBlocks)), dim3(kBlock),
                           0, ctx->stream, ctx->nC, ctx->cfo, ctx->cf, ctx->fgeo, ctx->q[ctx->cur], ctx->vol, ctx->dt,
                           ctx->diag, ctx->run, ctx->P.ratio_dtmax, ctx->K, ctx->diag32);
    else
        hipLaunchKernelGGL((dev::k_diag<MAXF>), dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, ctx->cfo,
                           ctx->cf, ctx->fgeo, ctx->q[ctx->cur], ctx->vol, ctx->dt, ctx->diag, nullptr,
                           ctx->P.ratio_dtmax, ctx->K, ctx->diag32);
    return check_launch(ctx, "k_diag");
}
bool launch_diag(mfs_ctx* ctx, bool clamp_only) {
    return ctx->maxf == 4 ? launch_diag_t<4>(ctx, clamp_only) : launch_diag_t<6>(ctx, clamp_only);
}

// Whether the sweeps run as two launches each (wide levels, thin tail).
bool sweep_split(const mfs_ctx* ctx) {
    return !ctx->sweep_by_tiles && ctx->tail_grid > 0 && ctx->tail_u < ctx->nchunk;
}

// Workgroups of the thin-tail launch of a split sweep.
int tail_workgroups(const mfs_ctx* ctx) {
    const int units = (ctx->nchunk - ctx->tail_u + ctx->sweep_waves - 1) / ctx->sweep_waves;
    return std::max(1, std::min({ctx->tail_grid, ctx->sweep_grid, units}));
}

// The two sweeps and the update.  `after` is enqueued after the forward
// sweep, or, with at_tail on a split sweep, after its wide levels, before
// its thin tail.
template <int LPC, typename R>
bool launch_sweeps_t(mfs_ctx* ctx, const std::function<bool()>& after, bool at_tail) {
    dev::SweepArgs<R> A;
    A.nC = ctx->nC;
    A.nq = ctx->nC + ctx->nB;
    A.nchunks = ctx->nchunk;
    A.ntiles = ctx->ntiles;
    A.tile_chunk = ctx->tile_chunk;
    A.pos = ctx->pos;
    A.flat = ctx->sweep_by_tiles ? 0 : 1;
    A.crun = ctx->crun;
    A.snb = ctx->snb;
    if constexpr (sizeof(R) == sizeof(double)) {
        A.sgeo = ctx->sgeo;
        A.q = ctx->q[ctx->cur];
        A.diag = ctx->diag;
        A.dqf = ctx->dqf;
        A.dqb = ctx->dq;
    } else {
        A.sgeo = ctx->sgeo32;
        A.q = ctx->q32;
        A.diag = ctx->diag32;
        A.dqf = ctx->dqf32;
        A.dqb = ctx->dq32;
    }
    A.q64 = ctx->q[ctx->cur];
    A.res = ctx->res;
    A.qout = ctx->q[1 - ctx->cur];
    A.err = ctx->err;
    A.depmask = ctx->depmask;
    A.cell0 = ctx->cell0;
    const dim3 block(64 * ctx->sweep_waves);
    const int units = A.flat ? (A.nchunks + ctx->sweep_waves - 1) / ctx->sweep_waves : A.ntiles;
    const dim3 grid((unsigned)std::max(1, std::min(ctx->sweep_grid, units)));
    // A split sweep runs its wide levels on the resident grid and its thin
    // tail on a small one: the forward sweep ends with the tail and the
    // backward sweep starts with it.
    const bool split = sweep_split(ctx);
    const int nt = A.nchunks - ctx->tail_u;  // steps of the tail
    const dim3 tgrid((unsigned)(split ? tail_workgroups(ctx) : 1));
    // Each launch takes its tickets from one counter set and zeroes the
    // other for the next launch.  The launches of an iteration alternate
    // between the two sets and are even in number, so every iteration's
    // first launch finds set 0 zeroed.
    int* const tk[2] = {ctx->tickets, ctx->tickets + kTicketGroups * kTicketStride};
    int set = 0;
    // ev0 / ev1: the timing events that this launch's start / end set
    auto launch = [&](const void* fn, dim3 g, int ubeg, int uend, const char* what, hipEvent_t ev0,
                      hipEvent_t ev1) {
        A.ubeg = ubeg;
        A.uend = uend;
        A.ngroups = std::min(kTicketGroups, (int)g.x);
        A.run_shift = g.x == grid.x ? dev::kWideRunShift : 0;
        A.tickets = tk[set];
        A.tickets_next = tk[1 - set];
        set ^= 1;
        void* args[] = {&A, &ctx->K};
        return hip_ok(ctx, hipExtLaunchKernel(fn, g, block, args, 0, ctx->stream, ev0, ev1, 0), what);
    };
    SweepTurn turn(ctx->device, ctx->stream);
    if (turn.error()) return sweep_turn_failed(ctx, turn.error());
    {
        SweepTimes t(ctx, split);
        const void* fwd = (const void*)dev::k_sweep<true, LPC, R>;
        if (split) {
            if (!launch(fwd, grid, 0, ctx->tail_u, "k_sweep<fwd>", t.start(0), t.stop(0))) return false;
            if (at_tail && after && !after()) return false;
            if (!launch(fwd, tgrid, ctx->tail_u, A.nchunks, "k_sweep<fwd> (thin tail)", t.start(1), t.stop(1)))
                return false;
        } else if (!launch(fwd, grid, 0, A.nchunks, "k_sweep<fwd>", t.start(0), t.stop(0))) {
            return false;
        }
        t.done("fwd_sweep", "fwd_sweep_wide", "fwd_sweep_tail");
    }
    if ((!at_tail || !split) && after && !after()) return false;
    {
        SweepTimes t(ctx, split);
        const void* bwd = (const void*)dev::k_sweep<false, LPC, R>;
        if (split) {
            if (!launch(bwd, tgrid, 0, nt, "k_sweep<bwd> (thin tail)", t.start(0), t.stop(0)) ||
                !launch(bwd, grid, nt, A.nchunks, "k_sweep<bwd>", t.start(1), t.stop(1)))
                return false;
        } else if (!launch(bwd, grid, 0, A.nchunks, "k_sweep<bwd>", t.start(0), t.stop(0))) {
            return false;
        }
        t.done("bwd_sweep", "bwd_sweep_tail", "bwd_sweep_wide");
    }
    if (!turn.done()) return sweep_turn_failed(ctx, turn.error());
    ctx->cur = 1 - ctx->cur;
    return true;
}
bool launch_sweeps(mfs_ctx* ctx, const std::function<bool()>& after = nullptr, bool at_tail = false) {
    if (ctx->precision == MFS_MIXED)
        return ctx->lpc == 4 ? launch_sweeps_t<4, float>(ctx, after, at_tail)
                             : launch_sweeps_t<8, float>(ctx, after, at_tail);
    return ctx->lpc == 4 ? launch_sweeps_t<4, double>(ctx, after, at_tail)
                         : launch_sweeps_t<8, double>(ctx, after, at_tail);
}

// CalculateGradient: node values, then the gradient.  With `next` set (the
// cfl of the next iteration in next_cfl), a tetrahedral context launches
// the gradient fused with the next iteration's residual / time-step pass.
template <int MAXF, typename R>
bool launch_gradient_t(mfs_ctx* ctx, bool next = false, double next_cfl = 0.0, hipStream_t st = nullptr,
                       const double* qs = nullptr) {
    if (!st) st = ctx->stream;
    const int n = ctx->nC + ctx->nB;
    const double* q = qs ? qs : ctx->q[ctx->cur];
    constexpr bool F64 = sizeof(R) == sizeof(double);
    R* qn;
    const R* ew;
    const typename dev::Real<R>::vec4* fgeo;
    R *gx, *gy, *gz;
    if constexpr (F64) {
        qn = ctx->qn;
        ew = ctx->ell_w;
        fgeo = ctx->fgeo;
        gx = ctx->gx;
        gy = ctx->gy;
        gz = ctx->gz;
    } else {
        qn = ctx->qn32;
        ew = ctx->ell_w32;
        fgeo = ctx->fgeo32;
        gx = ctx->gx32;
        gy = ctx->gy32;
        gz = ctx->gz32;
    }
    {
        Timed t(ctx, "node", st);
        hipLaunchKernelGGL(dev::k_node<R>, dim3(nblk_xcd(ctx->nN)), dim3(kBlock), 0, st, ctx->nN, ctx->nlo,
                           ctx->ell_off, ctx->ell_ab, ew, ctx->nden, q, qn);
        if (!check_launch(ctx, "k_node")) return false;
        if (ctx->nsym > 0) {
            hipLaunchKernelGGL(dev::k_node_symm<R>, dim3(nblk(ctx->nsym)), dim3(kBlock), 0, st, ctx->nsym,
                               ctx->sym_node, ctx->sym_n, qn);
            if (!check_launch(ctx, "k_node_symm")) return false;
        }
    }
    if (MAXF == 4 && ctx->snode && next) {
        Timed t(ctx, "grad_flux_ts", st);
        const typename dev::Real<R>::vec4* sgeo_r;
        if constexpr (F64)
            sgeo_r = ctx->sgeo;
        else
            sgeo_r = ctx->sgeo32;
        constexpr bool MIXED = !F64;
        hipLaunchKernelGGL((dev::k_grad_flux_ts<MIXED, R>), dim3(ctx->nblk_cell), dim3(kBlock), 0, st,
                           ctx->nC, n, ctx->snb, sgeo_r, ctx->snode, ctx->depmask, q, qn, ctx->vol, gx, gy, gz,
                           ctx->sgeo, ctx->sdn, ctx->sflip, ctx->res, ctx->dt, ctx->diag, ctx->part, ctx->K, next_cfl,
                           0.5 * next_cfl, MIXED ? ctx->q32 : nullptr, MIXED ? ctx->diag32 : nullptr);
        return check_launch(ctx, "k_grad_flux_ts");
    }
    {
        Timed t(ctx, "gradient", st);
        if (MAXF == 4 && ctx->snode) {
            const typename dev::Real<R>::vec4* sgeo;
            if constexpr (F64)
                sgeo = ctx->sgeo;
            else
                sgeo = ctx->sgeo32;
            // on the second stream: one workgroup per CU (a multiple of 8,
            // the same number on every XCD) claiming blocks; on the solver
            // stream the full grid
            const int gg = ((ctx->compute_units + dev::kXcd - 1) / dev::kXcd) * dev::kXcd;
            if (st != ctx->stream && gg > 0 && gg < ctx->nblk_cell)
                hipLaunchKernelGGL(dev::k_grad_tet_claim<R>, dim3(gg), dim3(kBlock), 0, st, ctx->nblk_cell, ctx->nC,
                                   n, ctx->snb, sgeo, ctx->snode, ctx->depmask, q, qn, ctx->vol, gx, gy, gz,
                                   ctx->tickets + kGradClaimOffset);
            else
                hipLaunchKernelGGL(dev::k_grad_tet<R>, dim3(ctx->nblk_cell), dim3(kBlock), 0, st, ctx->nC, n,
                                   ctx->snb, sgeo, ctx->snode, ctx->depmask, q, qn, ctx->vol, gx, gy, gz);
            if (!check_launch(ctx, "k_grad_tet")) return false;
        } else {
            hipLaunchKernelGGL((dev::k_grad<MAXF, R>), dim3(ctx->nblk_cell), dim3(kBlock), 0, st, ctx->nC,
                               n, ctx->cfo, ctx->cf, ctx->fcell, fgeo, ctx->fno, ctx->fn, ctx->btype, q, qn, ctx->vol,
                               gx, gy, gz);
            if (!check_launch(ctx, "k_grad")) return false;
        }
    }
    return true;
}
// SetGhostVariables on the current state (and its FP32 snapshot in mixed
// precision).
bool launch_ghosts(mfs_ctx* ctx) {
    if (ctx->nB == 0) return true;
    Timed t(ctx, "ghosts");
    const mfs_params& P = ctx->P;
    const dev::GhostConst G{P.rho_s, P.u_s, P.v_s, P.w_s, P.p_s};
    hipLaunchKernelGGL(dev::k_ghost, dim3(nblk(ctx->nB)), dim3(kBlock), 0, ctx->stream, ctx->nB, ctx->nC, ctx->gface,
                       ctx->fcell, ctx->fgeo, ctx->btype, ctx->K, G, ctx->q[ctx->cur],
                       ctx->precision == MFS_MIXED ? ctx->q32 : nullptr);
    return check_launch(ctx, "k_ghost");
}

bool launch_gradient(mfs_ctx* ctx, bool next = false, double next_cfl = 0.0, hipStream_t st = nullptr,
                     const double* qs = nullptr) {
    if (ctx->precision == MFS_MIXED)
        return ctx->maxf == 4 ? launch_gradient_t<4, float>(ctx, next, next_cfl, st, qs)
                              : launch_gradient_t<6, float>(ctx, next, next_cfl, st, qs);
    return ctx->maxf == 4 ? launch_gradient_t<4, double>(ctx, next, next_cfl, st, qs)
                          : launch_gradient_t<6, double>(ctx, next, next_cfl, st, qs);
}
// whether launch_gradient(ctx, true, ...) also runs the next residual pass
bool gradient_fuses_next(const mfs_ctx* ctx) { return ctx->maxf == 4 && ctx->snode; }

bool usable(mfs_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return false;
    }
    if (ctx->broken) {
        set_error("context is in an error state after an earlier failure");
        return false;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        set_error("hipSetDevice failed");
        return false;
    }
    return true;
}

void free_all(mfs_ctx* ctx) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs) (void)hipFree(b.p);
    ctx->bufs.clear();
    if (ctx->hist) (void)hipFree(ctx->hist);
    ctx->hist = nullptr;
    for (auto e : ctx->pending_events) (void)hipEventDestroy(e);
    ctx->pending.clear();
    ctx->pending_events.clear();
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    ctx->event_pool.clear();
    if (ctx->stream && ctx->owns_stream) {
        forget_sweep_stream(ctx->device, ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    ctx->stream = nullptr;
    if (ctx->gstream) (void)hipStreamDestroy(ctx->gstream);
    ctx->gstream = nullptr;
    for (hipEvent_t* e : {&ctx->sdone, &ctx->gdone[0], &ctx->gdone[1]})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
}

template <int LPC, typename R>
int sweep_occupancy(int* per_cu, int block) {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, dev::k_sweep<true, LPC, R>, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_sweep<false, LPC, R>, block, 0) != hipSuccess)
        return -1;
    *per_cu = std::min(a, b);
    return 0;
}

}  // namespace
}  // namespace mfs

// ===========================================================================
// C ABI: context functions.
using namespace mfs;

extern "C" int mfs_ctx_create(const mfs_mesh_view* mesh, const mfs_params* p, int precision, mfs_ctx** out) {
    return mfs_ctx_create_ex(mesh, p, precision, nullptr, out);
}

extern "C" int mfs_ctx_create_ex(const mfs_mesh_view* mesh, const mfs_params* p, int precision,
                                 const mfs_ctx_options* opt, mfs_ctx** out) {
    if (!mesh || !p || !out) {
        set_error("mfs_ctx_create: null argument");
        return MFS_ERR_ARG;
    }
    const int tile_cells = (opt && opt->tile_cells) ? opt->tile_cells : kDefaultTileCells;
    const int sweep_waves = (opt && opt->sweep_waves) ? opt->sweep_waves : 4;
    const int sweep_by_tiles = (opt && opt->sweep_by_tiles) ? 1 : 0;
    const int gradient_stream = (opt && opt->gradient_stream) ? opt->gradient_stream
                                : precision == MFS_MIXED   ? dev::kDefaultGradientStreamMixed
                                                           : dev::kDefaultGradientStreamFP64;
    if (tile_cells < 64 || (sweep_waves != 4 && sweep_waves != 8 && sweep_waves != 16)) {
        set_error("mfs_ctx_create: tile_cells must be >= 64 and sweep_waves 4, 8 or 16");
        return MFS_ERR_ARG;
    }
    if (gradient_stream < 1 || gradient_stream > 5) {
        set_error("mfs_ctx_create: gradient_stream must be 0 (default) or 1 to 5");
        return MFS_ERR_ARG;
    }
    const int tail_grid = (opt && opt->sweep_tail_grid) ? opt->sweep_tail_grid : dev::kDefaultTailGrid;
    if (tail_grid < -1) {
        set_error("mfs_ctx_create: sweep_tail_grid must be -1, 0 (default) or a workgroup count");
        return MFS_ERR_ARG;
    }
    *out = nullptr;
    if (precision != MFS_FP64 && precision != MFS_MIXED) {
        set_error("mfs_ctx_create: precision must be MFS_FP64 or MFS_MIXED");
        return MFS_ERR_ARG;
    }
    HostLayout L;
    std::string e = build_layout(*mesh, L, tile_cells);
    if (!e.empty()) {
        set_error("mfs_ctx_create: " + e);
        return MFS_ERR_MESH;
    }
    if (L.max_faces > dev::kMaxFaces) {
        set_error("mfs_ctx_create: cells with more than 6 faces are not supported");
        return MFS_ERR_MESH;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error("mfs_ctx_create: no HIP device available");
        return MFS_ERR_HIP;
    }
    mfs_ctx* ctx = new mfs_ctx();
    (void)hipGetDevice(&ctx->device);
    ctx->nC = L.nC;
    ctx->nF = L.nF;
    ctx->nB = L.nB;
    ctx->nN = L.nN;
    ctx->maxf = L.max_faces <= 4 ? 4 : 6;
    ctx->tile_cells = tile_cells;
    ctx->sweep_waves = sweep_waves;
    ctx->sweep_by_tiles = sweep_by_tiles;
    ctx->precision = precision;
    ctx->P = *p;
    ctx->K.gam = p->gam;
    ctx->K.gamm1 = p->gamm1;
    ctx->K.p_bar = p->p_bar;
    ctx->K.alf_l = p->alf_l;
    ctx->K.lhs_omga = p->lhs_omga;
    ctx->K.rho_min = p->rho_min;
    ctx->K.rho_max = p->rho_max;
    ctx->K.p_min = p->p_min;
    ctx->K.p_max = p->p_max;
    ctx->K.p_break = p->p_break;
    ctx->K.alpq = -0.2;
    ctx->K.phiq = 1. / 0.5;
    ctx->K.betq = 1.0 + ctx->K.alpq * ctx->K.phiq;
    ctx->nlev = L.nlev;
    ctx->nchunk = (int)(L.chunk.size() / 2);
    // the sweeps' thin tail: the thin levels that end the flat order (a
    // level's chunks are contiguous there and share the flag)
    ctx->tail_u = ctx->nchunk;
    while (ctx->tail_u > 0 && (L.crun[4 * (size_t)(ctx->tail_u - 1) + 1] & kThinChunkFlag)) --ctx->tail_u;
    if (ctx->tail_u == 0 || ctx->nchunk - ctx->tail_u < dev::kMinTailChunks) ctx->tail_u = ctx->nchunk;
    ctx->tail_cells = 0;
    for (int u = ctx->tail_u; u < ctx->nchunk; ++u) ctx->tail_cells += L.crun[4 * (size_t)u + 1] & 0xff;
    ctx->tail_grid = (tail_grid > 0 && !sweep_by_tiles && ctx->tail_u < ctx->nchunk) ? tail_grid : 0;
    ctx->ntiles = (int)L.tile_chunk.size() - 1;
    ctx->lpc = L.lpc;
    ctx->nblk_cell = (int)nblk_xcd(L.nC);
    ctx->node_of_dev.assign(L.node_of_dev.begin(), L.node_of_dev.end());
    ctx->cell_of_dev.assign(L.cell_of_dev.begin(), L.cell_of_dev.end());
    ctx->cell0 = L.dev_cell0;
    const size_t n = (size_t)L.nC + L.nB;
    ctx->gradient_stream = gradient_stream;
    bool ok = hip_ok(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (ok && gradient_stream >= 2)
        ok = hip_ok(ctx, hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking), "hipStreamCreate(gradient)") &&
             hip_ok(ctx, hipEventCreateWithFlags(&ctx->sdone, hipEventDisableTiming), "hipEventCreate") &&
             hip_ok(ctx, hipEventCreateWithFlags(&ctx->gdone[0], hipEventDisableTiming), "hipEventCreate") &&
             hip_ok(ctx, hipEventCreateWithFlags(&ctx->gdone[1], hipEventDisableTiming), "hipEventCreate");
    ok = ok && dalloc(ctx, ctx->q[0], 5 * n, "q0") && dalloc(ctx, ctx->q[1], 5 * n, "q1");
    ok = ok && dalloc(ctx, ctx->res, 5 * (size_t)L.nC, "res") && dalloc(ctx, ctx->dt, L.nC, "dt") &&
         dalloc(ctx, ctx->diag, L.nC, "diag") && dalloc(ctx, ctx->dq, 5 * n, "dq") &&
         dalloc(ctx, ctx->dqf, 5 * (size_t)L.nC, "dq_forward");
    ok = ok && dalloc(ctx, ctx->gx, 5 * n, "dqdx") && dalloc(ctx, ctx->gy, 5 * n, "dqdy") &&
         dalloc(ctx, ctx->gz, 5 * n, "dqdz") && dalloc(ctx, ctx->qn, 5 * (size_t)L.nN, "qnode");
    ok = ok && dalloc(ctx, ctx->fcell, L.nF, "fcell") && dalloc(ctx, ctx->fgeo, L.nF, "fgeo") &&
         dalloc(ctx, ctx->fdn, L.nF, "fdn") && dalloc(ctx, ctx->fno, L.nF + 1, "fno") &&
         dalloc(ctx, ctx->fn, L.fn.size(), "fn") && dalloc(ctx, ctx->btype, L.nB, "btype");
    if (ctx->maxf == 4) {
        // four faces per cell means tetrahedra: the gradient kernel reads
        // face f's three nodes at fn[3f..3f+2]
        for (int f = 0; f <= L.nF; ++f)
            if (L.fno[f] != 3 * f) {
                set_error("mfs_ctx_create: a mesh of 4-face cells has a face that is not a triangle");
                ok = false;
                break;
            }
    }
    ok = ok && dalloc(ctx, ctx->cfo, L.nC + 1, "cfo") && dalloc(ctx, ctx->cf, L.cf.size(), "cf") &&
         dalloc(ctx, ctx->vol, L.nC, "vol");
    // node entries, entry-major per wave of 64 nodes (padding: cell 0, skipped)
    const int nwave = (int)nblk_xcd(L.nN) * (kBlock / 64);
    std::vector<int> ell_off((size_t)nwave + 1, 0);
    for (int w = 0; w < nwave; ++w) {
        int K = 0;
        for (int p = 64 * w; p < std::min(L.nN, 64 * w + 64); ++p) K = std::max(K, L.nlo[p + 1] - L.nlo[p]);
        ell_off[w + 1] = ell_off[w] + 64 * K;
    }
    std::vector<int2> ell_ab((size_t)ell_off[nwave], make_int2(0, -1));
    std::vector<double> ell_w((size_t)ell_off[nwave], 0.0);
    for (int p = 0; p < L.nN; ++p)
        for (int k = 0; k < L.nlo[p + 1] - L.nlo[p]; ++k) {
            const size_t e = (size_t)ell_off[p / 64] + 64 * (size_t)k + (p % 64);
            ell_ab[e] = make_int2(L.nla[L.nlo[p] + k], L.nlb[L.nlo[p] + k]);
            ell_w[e] = L.nlw[L.nlo[p] + k];
        }
    ok = ok && dalloc(ctx, ctx->nlo, L.nN + 1, "nlo") && dalloc(ctx, ctx->ell_off, ell_off.size(), "ell_off") &&
         dalloc(ctx, ctx->ell_ab, ell_ab.size(), "ell_ab") && dalloc(ctx, ctx->ell_w, ell_w.size(), "ell_w") &&
         dalloc(ctx, ctx->nden, L.nN, "nden");
    ctx->nsym = (int)L.sym_node.size();
    ok = ok && dalloc(ctx, ctx->sym_node, L.sym_node.size(), "sym_node") &&
         dalloc(ctx, ctx->sym_n, L.sym_n.size(), "sym_n");
    ok = ok && dalloc(ctx, ctx->crun, (size_t)ctx->nchunk, "chunk records") &&
         dalloc(ctx, ctx->pos, (size_t)ctx->nchunk, "chunk positions") &&
         dalloc(ctx, ctx->tile_chunk, L.tile_chunk.size(), "tile_chunk") &&
         dalloc(ctx, ctx->snb, L.cfs.size(), "slot neighbours") &&
         dalloc(ctx, ctx->sgeo, L.cfs.size(), "slot geometry") && dalloc(ctx, ctx->depmask, L.nC, "depmask");
    const bool mixed = precision == MFS_MIXED;
    if (mixed) {
        ok = ok && dalloc(ctx, ctx->q32, 5 * n, "q32") && dalloc(ctx, ctx->diag32, L.nC, "diag32") &&
             dalloc(ctx, ctx->dqf32, 5 * (size_t)L.nC, "dq_forward32") && dalloc(ctx, ctx->dq32, 5 * n, "dq32") &&
             dalloc(ctx, ctx->gx32, 5 * n, "dqdx32") && dalloc(ctx, ctx->gy32, 5 * n, "dqdy32") &&
             dalloc(ctx, ctx->gz32, 5 * n, "dqdz32") && dalloc(ctx, ctx->qn32, 5 * (size_t)L.nN, "qnode32") &&
             dalloc(ctx, ctx->ell_w32, ell_w.size(), "ell_w32") && dalloc(ctx, ctx->fgeo32, L.nF, "fgeo32") &&
             dalloc(ctx, ctx->sgeo32, L.cfs.size(), "slot geometry (FP32)");
    }
    ok = ok && dalloc(ctx, ctx->part, partial_doubles((int)ctx->nblk_cell), "partials") && dalloc(ctx, ctx->run, 8, "run") &&
         dalloc(ctx, ctx->err, 4, "err") &&
         dalloc(ctx, ctx->tickets, kTicketInts, "sweep tickets") &&
         dalloc(ctx, ctx->gface, (size_t)L.nB, "ghost faces");
    if (L.tri_slots) ok = ok && dalloc(ctx, ctx->snode, 12 * (size_t)L.nC, "slot nodes");
    ok = ok && dalloc(ctx, ctx->sdn, L.cfs.size(), "slot distances") && dalloc(ctx, ctx->sflip, L.nC, "slot sides");
    ok = ok && ensure_hist(ctx, 64);
    // per-device set-up: the sweep order's event and the address of the
    // static scratch of mfs_step_noalloc on this device
    ok = ok && sweep_order_setup(ctx->device) && noalloc_scratch_setup(ctx->device);
    if (ok) {
        // resident workgroups of the sweep launch
        int per_cu = 0, cus = 0;
        const bool mixed = ctx->precision == MFS_MIXED;
        const int blk = 64 * ctx->sweep_waves;
        const int occ = ctx->lpc == 4
                            ? (mixed ? sweep_occupancy<4, float>(&per_cu, blk) : sweep_occupancy<4, double>(&per_cu, blk))
                            : (mixed ? sweep_occupancy<8, float>(&per_cu, blk) : sweep_occupancy<8, double>(&per_cu, blk));
        ok = occ == 0 &&
             hip_ok(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
        if (!ok) set_error("mfs_ctx_create: occupancy query failed");
        // co-resident workgroups
        ctx->sweep_grid = std::max(1, std::min(per_cu, kSweepWavesPerCU / ctx->sweep_waves) * cus);
        ctx->compute_units = cus;
    }
    if (ok) {
        std::vector<int2> fc(L.nF);
        std::vector<double2> fd(L.nF);
        std::vector<int> gf(std::max(L.nB, 1), 0);
        for (int i = 0; i < L.nF; ++i) {
            fc[i] = make_int2(L.c1[i], L.c2[i]);
            fd[i] = make_double2(L.dn1[i], L.dn2[i]);
            if (L.c2[i] >= L.nC) gf[L.c2[i] - L.nC] = i;
        }
        std::vector<double> qa(5 * n);
        for (size_t c = 0; c < n; ++c)
            for (int j = 0; j < 5; ++j) qa[c * 5 + j] = L.q[(size_t)j * n + c];
        ok = upload(ctx, ctx->q[0], qa.data(), 5 * n, "upload q") && upload(ctx, ctx->q[1], qa.data(), 5 * n, "upload q") &&
             upload(ctx, ctx->fcell, fc.data(), L.nF, "upload fcell") &&
             upload(ctx, ctx->fgeo, reinterpret_cast<const double4*>(L.fgeo.data()), L.nF, "upload fgeo") &&
             upload(ctx, ctx->fdn, fd.data(), L.nF, "upload fdn") && upload(ctx, ctx->fno, L.fno.data(), L.nF + 1, "upload fno") &&
             upload(ctx, ctx->fn, L.fn.data(), L.fn.size(), "upload fn") &&
             upload(ctx, ctx->btype, L.btype.data(), L.nB, "upload btype") &&
             upload(ctx, ctx->cfo, L.cfo.data(), L.nC + 1, "upload cfo") && upload(ctx, ctx->cf, L.cf.data(), L.cf.size(), "upload cf") &&
             upload(ctx, ctx->vol, L.vol.data(), L.nC, "upload vol") && upload(ctx, ctx->nlo, L.nlo.data(), L.nN + 1, "upload nlo") &&
             upload(ctx, ctx->ell_off, ell_off.data(), ell_off.size(), "upload ell_off") &&
             upload(ctx, ctx->ell_ab, ell_ab.data(), ell_ab.size(), "upload ell_ab") &&
             upload(ctx, ctx->ell_w, ell_w.data(), ell_w.size(), "upload ell_w") &&
             upload(ctx, ctx->nden, L.nden.data(), L.nN, "upload nden") &&
             upload(ctx, ctx->sym_node, L.sym_node.data(), L.sym_node.size(), "upload sym") &&
             upload(ctx, ctx->sym_n, L.sym_n.data(), L.sym_n.size(), "upload sym_n") &&
             upload(ctx, ctx->crun, reinterpret_cast<const int4*>(L.crun.data()), (size_t)ctx->nchunk, "upload crun") &&
             upload(ctx, ctx->pos, L.pos.data(), L.pos.size(), "upload pos") &&
             upload(ctx, ctx->tile_chunk, L.tile_chunk.data(), L.tile_chunk.size(), "upload tile_chunk") &&
             upload(ctx, ctx->depmask, L.depmask.data(), L.nC, "upload depmask") &&
             upload(ctx, ctx->gface, gf.data(), (size_t)L.nB, "upload gface");
        // the slot tables, built on the device from the slot face ids
        int* dcfs = nullptr;
        ok = ok && hip_ok(ctx, hipMalloc((void**)&dcfs, L.cfs.size() * sizeof(int)), "hipMalloc(slot faces)");
        if (ok) {
            ok = upload(ctx, dcfs, L.cfs.data(), L.cfs.size(), "upload slot faces");
            if (ok) {
                const unsigned gC = (unsigned)((L.nC + kBlock - 1) / kBlock);
                auto kern = L.lpc == 4 ? dev::k_slot_tables<4> : dev::k_slot_tables<8>;
                hipLaunchKernelGGL(kern, dim3(gC), dim3(kBlock), 0, ctx->stream, L.nC, dcfs, ctx->cfo, ctx->fcell,
                                   ctx->fgeo, ctx->fdn, ctx->fno, ctx->fn, ctx->snb, ctx->sgeo,
                                   mixed ? ctx->sgeo32 : nullptr, ctx->sdn, ctx->sflip, ctx->snode);
                ok = check_launch(ctx, "k_slot_tables") &&
                     hip_ok(ctx, hipStreamSynchronize(ctx->stream), "slot tables");
            }
            (void)hipFree(dcfs);
        }
        const double run0[8] = {0.0, kBIGd, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        ok = ok && upload(ctx, ctx->run, run0, 8, "upload run");
        std::vector<float> f32;
        std::vector<float4> g32;
        if (ok && mixed) {
            // FP32 copies of the operator's static inputs; the ghost part of
            // the state snapshot is fixed (ghosts are frozen after set-up)
            g32.resize(L.nF);
            for (int i = 0; i < L.nF; ++i)
                g32[i] = make_float4((float)L.fgeo[4 * (size_t)i], (float)L.fgeo[4 * (size_t)i + 1],
                                     (float)L.fgeo[4 * (size_t)i + 2], (float)L.fgeo[4 * (size_t)i + 3]);
            f32.resize(std::max(ell_w.size(), 5 * n));
            for (size_t k = 0; k < ell_w.size(); ++k) f32[k] = (float)ell_w[k];
            ok = upload(ctx, ctx->fgeo32, g32.data(), L.nF, "upload fgeo32") &&
                 upload(ctx, ctx->ell_w32, f32.data(), ell_w.size(), "upload ell_w32");
            for (size_t k = 0; k < 5 * n; ++k) f32[k] = (float)qa[k];
            ok = ok && upload(ctx, ctx->q32, f32.data(), 5 * n, "upload q32");
            // hipMemcpyAsync reads the host buffers later: wait before they go out of scope
            ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "upload sync");
        }
        if (ok) {
            double sentinel;
            std::memcpy(&sentinel, &dev::kSentinel, sizeof(sentinel));
            const size_t cnt = 5 * (size_t)L.nC;
            hipLaunchKernelGGL(dev::k_fill, dim3(nblk(cnt)), dim3(kBlock), 0, ctx->stream, ctx->dqf, cnt, sentinel);
            ok = check_launch(ctx, "k_fill(dq_forward)");
            if (ok && mixed) {
                float sent32;
                const unsigned bits = dev::Real<float>::kSent;
                std::memcpy(&sent32, &bits, sizeof(sent32));
                hipLaunchKernelGGL(dev::k_fill_f, dim3(nblk(cnt)), dim3(kBlock), 0, ctx->stream, ctx->dqf32, cnt,
                                   sent32);
                ok = check_launch(ctx, "k_fill(dq_forward32)") &&
                     hip_ok(ctx, hipMemsetAsync(ctx->dq32, 0, sizeof(float) * 5 * n, ctx->stream), "memset dq32") &&
                     hip_ok(ctx, hipMemsetAsync(ctx->gx32, 0, sizeof(float) * 5 * n, ctx->stream), "memset dqdx32") &&
                     hip_ok(ctx, hipMemsetAsync(ctx->gy32, 0, sizeof(float) * 5 * n, ctx->stream), "memset dqdy32") &&
                     hip_ok(ctx, hipMemsetAsync(ctx->gz32, 0, sizeof(float) * 5 * n, ctx->stream), "memset dqdz32");
            }
        }
        ok = ok && hip_ok(ctx, hipMemsetAsync(ctx->err, 0, sizeof(int) * 4, ctx->stream), "memset err") &&
             hip_ok(ctx,
                    hipMemsetAsync(ctx->tickets, 0, sizeof(int) * kTicketInts, ctx->stream),
                    "memset tickets") &&
             hip_ok(ctx, hipMemsetAsync(ctx->dq, 0, sizeof(double) * 5 * n, ctx->stream), "memset dq") &&
             hip_ok(ctx, hipMemsetAsync(ctx->res, 0, sizeof(double) * 5 * L.nC, ctx->stream), "memset res") &&
             hip_ok(ctx, hipMemsetAsync(ctx->dt, 0, sizeof(double) * L.nC, ctx->stream), "memset dt") &&
             hip_ok(ctx, hipMemsetAsync(ctx->diag, 0, sizeof(double) * L.nC, ctx->stream), "memset diag") &&
             hip_ok(ctx, hipMemsetAsync(ctx->gx, 0, sizeof(double) * 5 * n, ctx->stream), "memset dqdx") &&
             hip_ok(ctx, hipMemsetAsync(ctx->gy, 0, sizeof(double) * 5 * n, ctx->stream), "memset dqdy") &&
             hip_ok(ctx, hipMemsetAsync(ctx->gz, 0, sizeof(double) * 5 * n, ctx->stream), "memset dqdz");
        // AllocAndCalQuantityGradient (cal_gradient.cpp:50-78): gradients of the initial state.
        ok = ok && launch_gradient(ctx);
        ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "create sync");
    }
    if (!ok) {
        free_all(ctx);
        delete ctx;
        return MFS_ERR_HIP;
    }
    *out = ctx;
    return MFS_OK;
}

extern "C" int mfs_ctx_destroy(mfs_ctx* ctx) {
    if (!ctx) return MFS_OK;
    (void)hipSetDevice(ctx->device);
    free_all(ctx);
    delete ctx;
    return MFS_OK;
}

extern "C" int mfs_zero_residuals(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    const size_t cnt = 5 * (size_t)ctx->nC;
    hipLaunchKernelGGL(dev::k_fill, dim3(nblk(cnt)), dim3(kBlock), 0, ctx->stream, ctx->res, cnt, 0.0);
    return check_launch(ctx, "k_fill(res)") ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_inviscid_flux(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (!launch_flux(ctx, true)) return MFS_ERR_HIP;
    return launch_reduce(ctx, 1, -1) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_compute_timestep(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    const double cfl = cfl_of_iter(ctx->P, ctx->iter_done);
    const double cfl_min = 0.5 * cfl;
    ctx->cfl_min = cfl_min;
    if (!launch_timestep(ctx, false, cfl, cfl_min)) return MFS_ERR_HIP;
    if (!launch_reduce(ctx, 2, -1)) return MFS_ERR_HIP;
    // ratio_dtmax clamp (compute_timestep.cpp:929-945)
    return launch_diag(ctx, true) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_forward_lusgs(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (!launch_diag(ctx, false)) return MFS_ERR_HIP;
    if (ctx->precision == MFS_MIXED && !launch_mixed_inputs(ctx)) return MFS_ERR_HIP;
    return launch_sweeps(ctx) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_pass_interface_data(mfs_ctx* ctx) { return usable(ctx) ? MFS_OK : MFS_ERR_STATE; }

extern "C" int mfs_set_ghost_variables(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    return launch_ghosts(ctx) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_ctx_set_ghost_refresh(mfs_ctx* ctx, int enable) {
    if (!ctx) {
        set_error("mfs_ctx_set_ghost_refresh: null context");
        return MFS_ERR_ARG;
    }
    ctx->ghost_refresh = enable ? 1 : 0;
    return MFS_OK;
}

extern "C" int mfs_calculate_gradient(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    return launch_gradient(ctx) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_iteration_done(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (!ensure_hist(ctx, ctx->hist_n + 1)) return MFS_ERR_HIP;
    if (!launch_reduce(ctx, 4, ctx->hist_n)) return MFS_ERR_HIP;
    ctx->hist_n++;
    ctx->iter_done++;
    return MFS_OK;
}

namespace {
// The iterations of mfs_step; the history records go to ctx->hist from
// slot ctx->hist_n on, which must hold nsteps more records.
int step_iterations(mfs_ctx* ctx, int nsteps) {
    // The gradient feeds nothing in the iteration (CalculateGradient's
    // output is not read by the flux, time step or sweeps), so it may run
    // beside the next iteration; only the state buffer it reads must
    // outlive it.  gradient_stream 2/3/4/5: iteration k's node values and
    // gradient start on the second stream after iteration k's update /
    // iteration k+1's residual pass / iteration k+1's forward sweep / the
    // wide levels of that sweep (before its thin tail).
    const int mode = ctx->gstream ? ctx->gradient_stream : 1;
    const bool side = mode >= 2;
    const bool fuse = !side && gradient_fuses_next(ctx);
    const double* grad_q = nullptr;  // state of the gradient not yet launched
    int grad_par = 0;
    // records `ev` on the solver stream and starts the pending gradient
    // behind it on the second stream
    auto side_gradient = [&](hipEvent_t ev) -> bool {
        if (!grad_q) return true;
        if (!hip_ok(ctx, hipEventRecord(ev, ctx->stream), "record gradient start") ||
            !hip_ok(ctx, hipStreamWaitEvent(ctx->gstream, ev, 0), "gradient waits") ||
            !launch_gradient(ctx, false, 0.0, ctx->gstream, grad_q) ||
            !hip_ok(ctx, hipEventRecord(ctx->gdone[grad_par], ctx->gstream), "record gradient"))
            return false;
        ctx->gpend[grad_par] = true;
        grad_q = nullptr;
        return true;
    };
    for (int s = 0; s < nsteps; ++s) {
        const double cfl = cfl_of_iter(ctx->P, ctx->iter_done);
        const double cfl_min = 0.5 * cfl;
        ctx->cfl_min = cfl_min;
        // ZeroResiduals + InviscidFlux + ComputeTimeStep with the LU-SGS
        // diagonal, in one pass over the faces (res written from zero);
        // after the first iteration the previous gradient launch did it
        // when fused.
        if ((s == 0 || !fuse) && !launch_flux_ts(ctx, cfl, cfl_min)) return MFS_ERR_HIP;
        if (side && mode == 3 && !side_gradient(ctx->sdone)) return MFS_ERR_HIP;
        if (!launch_reduce(ctx, 1 | 2 | 4, ctx->hist_n)) return MFS_ERR_HIP;
        if (!launch_diag(ctx, true)) return MFS_ERR_HIP;  // acts only when the ratio_dtmax clamp triggers
        // The backward sweep writes the state buffer that the gradient of
        // two iterations back reads.  (Waiting between the forward sweep's
        // two launches instead, beside the event that starts the next
        // gradient, let the node values reach the CUs before the thin tail
        // and slowed the tail by 90 us.)
        const int par = ctx->iter_done & 1;
        if (side && ctx->gpend[par]) {
            if (!hip_ok(ctx, hipStreamWaitEvent(ctx->stream, ctx->gdone[par], 0), "wait gradient")) return MFS_ERR_HIP;
            ctx->gpend[par] = false;
        }
        // ForwardLUSGS: sweeps + update
        std::function<bool()> after_fwd;
        if (side && mode >= 4 && grad_q) after_fwd = [&] { return side_gradient(ctx->sdone); };
        if (!launch_sweeps(ctx, after_fwd, mode == 5)) return MFS_ERR_HIP;
        // SetGhostVariables, when enabled (main.cpp:283 has the call commented out)
        if (ctx->ghost_refresh && !launch_ghosts(ctx)) return MFS_ERR_HIP;
        // CalculateGradient (fused: with the next iteration's first pass)
        if (side) {
            grad_q = ctx->q[ctx->cur];
            grad_par = par;
            if ((mode == 2 || s + 1 == nsteps) && !side_gradient(ctx->sdone)) return MFS_ERR_HIP;
        } else {
            const bool next = fuse && s + 1 < nsteps;
            if (!launch_gradient(ctx, next, next ? cfl_of_iter(ctx->P, ctx->iter_done + 1) : 0.0)) return MFS_ERR_HIP;
        }
        ctx->hist_n++;
        ctx->iter_done++;
    }
    if (side && nsteps > 0) {
        // join: everything later on the solver stream (entry points,
        // downloads, mfs_sync) comes after the last gradient
        const int last = (ctx->iter_done - 1) & 1;
        if (!hip_ok(ctx, hipStreamWaitEvent(ctx->stream, ctx->gdone[last], 0), "join gradient")) return MFS_ERR_HIP;
        ctx->gpend[0] = ctx->gpend[1] = false;
    }
    return MFS_OK;
}

// For the duration of one mfs_step_noalloc call: the context's partials,
// ticket counters and history point into the static scratch and the
// caller's record buffer, and kernel timing (which creates events) is off.
struct NoAllocView {
    mfs_ctx* ctx;
    double* part;
    int* tickets;
    double* hist;
    int hist_n, hist_cap;
    bool timing;
    NoAllocView(mfs_ctx* c, double* sp, int* st, double* h, int n)
        : ctx(c), part(c->part), tickets(c->tickets), hist(c->hist), hist_n(c->hist_n), hist_cap(c->hist_cap),
          timing(c->timing) {
        c->part = sp;
        c->tickets = st;
        c->hist = h;
        c->hist_n = 0;
        c->hist_cap = n;
        c->timing = 0;
    }
    ~NoAllocView() {
        ctx->part = part;
        ctx->tickets = tickets;
        ctx->hist = hist;
        ctx->hist_n = hist_n;
        ctx->hist_cap = hist_cap;
        ctx->timing = timing;
    }
};
}  // namespace

extern "C" int mfs_step(mfs_ctx* ctx, int nsteps) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (nsteps < 0) {
        set_error("mfs_step: negative step count");
        return MFS_ERR_ARG;
    }
    if (!ensure_hist(ctx, ctx->hist_n + nsteps)) return MFS_ERR_HIP;
    return step_iterations(ctx, nsteps);
}

extern "C" int mfs_step_noalloc(mfs_ctx* ctx, int nsteps, double* hist_dev) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (nsteps < 0 || (nsteps > 0 && !hist_dev)) {
        set_error("mfs_step_noalloc: negative step count or null history buffer");
        return MFS_ERR_ARG;
    }
    if (!ctx->owns_stream) {
        set_error("mfs_step_noalloc: the context belongs to a batch; step it with mfs_batch_step");
        return MFS_ERR_STATE;
    }
    if (ctx->nC > kNoAllocMaxCells || partial_doubles((int)ctx->nblk_cell) > kNoAllocPartDoubles) {
        set_error("mfs_step_noalloc: " + std::to_string(ctx->nC) + " cells exceed the static scratch's limit of " +
                  std::to_string(kNoAllocMaxCells) + " cells (use mfs_step)");
        return MFS_ERR_ARG;
    }
    char* sp = static_cast<char*>(noalloc_scratch(ctx->device));
    if (!sp) {
        set_error("mfs_step_noalloc: the static scratch of this device is not set up");
        return MFS_ERR_STATE;
    }
    double* part = reinterpret_cast<double*>(sp);
    int* tickets = reinterpret_cast<int*>(sp + kNoAllocPartDoubles * sizeof(double));
    // The counters start from zero (an earlier call may have ended with the
    // backward ones in use); a fill kernel rather than a memset node, so the
    // call can be captured in a graph.  The partials need no clearing: each
    // residual pass writes every block's record before the reduction reads it.
    hipLaunchKernelGGL(dev::k_fill_i, dim3(nblk(kNoAllocTicketInts)), dim3(kBlock), 0, ctx->stream, tickets,
                       kNoAllocTicketInts, 0);
    if (!check_launch(ctx, "k_fill_i(scratch tickets)")) return MFS_ERR_HIP;
    NoAllocView view(ctx, part, tickets, hist_dev, nsteps);
    return step_iterations(ctx, nsteps);
}

extern "C" int mfs_sync(mfs_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return MFS_ERR_ARG;
    }
    (void)hipSetDevice(ctx->device);
    if (!hip_ok(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) return MFS_ERR_HIP;
    int err[4] = {0, 0, 0, 0};
    if (!hip_ok(ctx, hipMemcpy(err, ctx->err, sizeof(err), hipMemcpyDeviceToHost), "read error flags")) return MFS_ERR_HIP;
    if (ctx->timing) collect_timing(ctx);
    if (err[0]) {
        ctx->broken = true;
        set_error("LU-SGS sweep exceeded its dependency wait limit");
        return MFS_ERR_STATE;
    }
    if (ctx->broken) {
        set_error("context is in an error state after an earlier failure");
        return MFS_ERR_STATE;
    }
    return MFS_OK;
}

extern "C" int mfs_history_count(mfs_ctx* ctx) { return ctx ? ctx->hist_n : 0; }

extern "C" int mfs_get_history(mfs_ctx* ctx, double* out, int max_records) {
    if (!ctx || !out || max_records < 0) {
        set_error("mfs_get_history: bad argument");
        return MFS_ERR_ARG;
    }
    int r = mfs_sync(ctx);
    if (r != MFS_OK) return r;
    const int cnt = std::min(max_records, ctx->hist_n);
    if (cnt > 0 && !hip_ok(ctx, hipMemcpy(out, ctx->hist, sizeof(double) * 7 * cnt, hipMemcpyDeviceToHost), "read history"))
        return MFS_ERR_HIP;
    return cnt;
}

// Field geometry: reference-layout element count and the device array.
namespace {
enum class FieldKind { StateBlock, CellMajor5, Scalar, SoA, NodeMajor5 };
struct FieldRef {
    double* dev = nullptr;
    size_t count = 0;   // reference-layout element count
    size_t rows = 0;    // cells or nodes for cell/node-major arrays
    FieldKind kind = FieldKind::Scalar;
    int block = 0;      // state variable for StateBlock
    float* dev32 = nullptr;  // FP32 array of a mixed-precision context (read-only fields)
};
FieldRef field_ref(mfs_ctx* ctx, int field) {
    const size_t n = (size_t)ctx->nC + ctx->nB;
    FieldRef f;
    switch (field) {
        case MFS_F_RHO:
        case MFS_F_U:
        case MFS_F_V:
        case MFS_F_W:
        case MFS_F_P:
            f = {ctx->q[ctx->cur], n, n, FieldKind::StateBlock, field};
            break;
        case MFS_F_RES: f = {ctx->res, 5 * (size_t)ctx->nC, (size_t)ctx->nC, FieldKind::CellMajor5, 0}; break;
        case MFS_F_DT: f = {ctx->dt, (size_t)ctx->nC, (size_t)ctx->nC, FieldKind::Scalar, 0}; break;
        case MFS_F_DQ: f = {ctx->dq, 5 * n, n, FieldKind::CellMajor5, 0}; break;
        case MFS_F_DQDX: f = {ctx->gx, 5 * n, n, FieldKind::SoA, 0}; break;
        case MFS_F_DQDY: f = {ctx->gy, 5 * n, n, FieldKind::SoA, 0}; break;
        case MFS_F_DQDZ: f = {ctx->gz, 5 * n, n, FieldKind::SoA, 0}; break;
        case MFS_F_DIAG: f = {ctx->diag, (size_t)ctx->nC, (size_t)ctx->nC, FieldKind::Scalar, 0}; break;
        case MFS_F_QNODE: f = {ctx->qn, 5 * (size_t)ctx->nN, (size_t)ctx->nN, FieldKind::NodeMajor5, 0}; break;
        default: break;
    }
    if (ctx->precision == MFS_MIXED) {
        // the implicit operator's results and the gradients live in FP32
        switch (field) {
            case MFS_F_DQ: f.dev32 = ctx->dq32; break;
            case MFS_F_DQDX: f.dev32 = ctx->gx32; break;
            case MFS_F_DQDY: f.dev32 = ctx->gy32; break;
            case MFS_F_DQDZ: f.dev32 = ctx->gz32; break;
            case MFS_F_QNODE: f.dev32 = ctx->qn32; break;
            default: break;
        }
    }
    return f;
}
}  // namespace

extern "C" size_t mfs_field_count(mfs_ctx* ctx, int field) { return ctx ? field_ref(ctx, field).count : 0; }

extern "C" int mfs_download(mfs_ctx* ctx, int field, double* dst, size_t count) {
    if (!ctx || !dst) {
        set_error("mfs_download: null argument");
        return MFS_ERR_ARG;
    }
    FieldRef f = field_ref(ctx, field);
    if (!f.dev || count != f.count) {
        set_error("mfs_download: unknown field or wrong count");
        return MFS_ERR_ARG;
    }
    int r = mfs_sync(ctx);
    if (r != MFS_OK) return r;
    std::vector<double> tmp(f.kind == FieldKind::Scalar ? f.rows : 5 * f.rows);
    if (f.dev32) {
        std::vector<float> t32(tmp.size());
        if (!hip_ok(ctx, hipMemcpy(t32.data(), f.dev32, sizeof(float) * t32.size(), hipMemcpyDeviceToHost), "download"))
            return MFS_ERR_HIP;
        for (size_t k = 0; k < t32.size(); ++k) tmp[k] = (double)t32[k];
    } else if (!hip_ok(ctx, hipMemcpy(tmp.data(), f.dev, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost),
                       "download")) {
        return MFS_ERR_HIP;
    }
    // device row -> mesh row: cells through cell_of_dev, ghosts unchanged, nodes through node_of_dev
    const size_t nC = (size_t)ctx->nC;
    auto mesh_row = [&](size_t d) -> size_t {
        if (f.kind == FieldKind::NodeMajor5) return (size_t)ctx->node_of_dev[d];
        return d < nC ? (size_t)ctx->cell_of_dev[d] : d;
    };
    switch (f.kind) {
        case FieldKind::Scalar:
            for (size_t d = 0; d < f.rows; ++d) dst[mesh_row(d)] = tmp[d];
            break;
        case FieldKind::SoA:
            for (int j = 0; j < 5; ++j)
                for (size_t d = 0; d < f.rows; ++d) dst[(size_t)j * f.rows + mesh_row(d)] = tmp[(size_t)j * f.rows + d];
            break;
        case FieldKind::StateBlock:
            for (size_t d = 0; d < f.rows; ++d) dst[mesh_row(d)] = tmp[d * 5 + f.block];
            break;
        case FieldKind::CellMajor5:
        case FieldKind::NodeMajor5:
            for (size_t d = 0; d < f.rows; ++d)
                for (int j = 0; j < 5; ++j) dst[(size_t)j * f.rows + mesh_row(d)] = tmp[d * 5 + j];
            break;
    }
    return MFS_OK;
}

extern "C" int mfs_upload(mfs_ctx* ctx, int field, const double* src, size_t count) {
    if (!usable(ctx) || !src) return MFS_ERR_ARG;
    FieldRef f = field_ref(ctx, field);
    const bool allowed = f.kind == FieldKind::StateBlock || field == MFS_F_RES || field == MFS_F_DT;
    if (!f.dev || count != f.count || !allowed) {
        set_error("mfs_upload: unknown or read-only field, or wrong count");
        return MFS_ERR_ARG;
    }
    if (!hip_ok(ctx, hipStreamSynchronize(ctx->stream), "upload sync")) return MFS_ERR_HIP;
    const size_t nC = (size_t)ctx->nC;
    auto mesh_row = [&](size_t d) -> size_t { return d < nC ? (size_t)ctx->cell_of_dev[d] : d; };
    std::vector<double> tmp(f.kind == FieldKind::Scalar ? f.rows : 5 * f.rows);
    if (f.kind == FieldKind::Scalar) {
        for (size_t d = 0; d < f.rows; ++d) tmp[d] = src[mesh_row(d)];
    } else if (f.kind == FieldKind::StateBlock) {
        if (!hip_ok(ctx, hipMemcpy(tmp.data(), f.dev, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost), "upload read"))
            return MFS_ERR_HIP;
        for (size_t d = 0; d < f.rows; ++d) tmp[d * 5 + f.block] = src[mesh_row(d)];
        // both state buffers, so the ghost values stay identical in each
        bool ok = hip_ok(ctx, hipMemcpy(ctx->q[0], tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice),
                         "upload") &&
                  hip_ok(ctx, hipMemcpy(ctx->q[1], tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice),
                         "upload");
        if (ok && ctx->precision == MFS_MIXED) {  // the FP32 snapshot, ghosts included
            std::vector<float> t32(tmp.size());
            for (size_t k = 0; k < tmp.size(); ++k) t32[k] = (float)tmp[k];
            ok = hip_ok(ctx, hipMemcpy(ctx->q32, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice),
                        "upload");
        }
        return ok ? MFS_OK : MFS_ERR_HIP;
    } else {
        for (size_t d = 0; d < f.rows; ++d)
            for (int j = 0; j < 5; ++j) tmp[d * 5 + j] = src[(size_t)j * f.rows + mesh_row(d)];
    }
    return hip_ok(ctx, hipMemcpy(f.dev, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice), "upload")
               ? MFS_OK
               : MFS_ERR_HIP;
}

extern "C" int mfs_get_scalars(mfs_ctx* ctx, mfs_scalars* s) {
    if (!ctx || !s) {
        set_error("mfs_get_scalars: null argument");
        return MFS_ERR_ARG;
    }
    int r = mfs_sync(ctx);
    if (r != MFS_OK) return r;
    double run[8];
    if (!hip_ok(ctx, hipMemcpy(run, ctx->run, sizeof(run), hipMemcpyDeviceToHost), "read scalars")) return MFS_ERR_HIP;
    s->iter_done = ctx->iter_done;
    s->dt_max = run[0];
    s->dt_min = run[1];
    s->cfl_min = ctx->cfl_min;
    return MFS_OK;
}

extern "C" int mfs_ctx_get_info(mfs_ctx* ctx, mfs_ctx_info* info) {
    if (!ctx || !info) {
        set_error("mfs_ctx_get_info: null argument");
        return MFS_ERR_ARG;
    }
    info->nTCell = ctx->nC;
    info->nTFace = ctx->nF;
    info->nBFace = ctx->nB;
    info->nTNode = ctx->nN;
    info->precision = ctx->precision;
    info->sweep_levels = ctx->nlev;
    info->device_bytes = ctx->device_bytes;
    info->tile_cells = ctx->tile_cells;
    info->sweep_waves = ctx->sweep_waves;
    info->sweep_tiles = ctx->ntiles;
    info->sweep_workgroups = ctx->sweep_by_tiles ? std::min(ctx->sweep_grid, ctx->ntiles)
                                                 : std::min(ctx->sweep_grid, (ctx->nchunk + ctx->sweep_waves - 1) /
                                                                                 ctx->sweep_waves);
    info->sweep_by_tiles = ctx->sweep_by_tiles;
    info->compute_units = ctx->compute_units;
    info->sweep_tail_chunks = ctx->tail_grid > 0 ? ctx->nchunk - ctx->tail_u : 0;
    info->sweep_tail_cells = ctx->tail_grid > 0 ? ctx->tail_cells : 0;
    info->sweep_tail_workgroups = ctx->tail_grid > 0 ? tail_workgroups(ctx) : 0;
    return MFS_OK;
}

extern "C" int mfs_ctx_enable_timing(mfs_ctx* ctx, int enable) {
    if (!ctx) return MFS_ERR_ARG;
    ctx->timing = enable == 2 ? 2 : (enable ? 1 : 0);
    return MFS_OK;
}

extern "C" int mfs_ctx_get_timing(mfs_ctx* ctx, const char* name, double* ms, int* launches) {
    if (!ctx || !name || !ms || !launches) return MFS_ERR_ARG;
    int r = mfs_sync(ctx);
    if (r != MFS_OK) return r;
    auto it = ctx->timers.find(name);
    *ms = it == ctx->timers.end() ? 0.0 : it->second.ms;
    *launches = it == ctx->timers.end() ? 0 : it->second.launches;
    return MFS_OK;
}

extern "C" int mfs_ctx_reset_timing(mfs_ctx* ctx) {
    if (!ctx) return MFS_ERR_ARG;
    int r = mfs_sync(ctx);
    ctx->timers.clear();
    return r;
}

// ===========================================================================
// Batched multi-solve (mfs_batch_*): several solves stepped together, one
// launch per kernel for all of them, on the batch's stream.
struct mfs_batch {
    int device = 0;
    int precision = MFS_FP64;
    int maxf = 0, lpc = 0;
    hipStream_t stream = nullptr;
    std::vector<mfs_ctx*> slots;  // removed solves leave nullptr
    bool dirty = true;            // block maps and sweep items need a rebuild
    // device-side maps for the current set of solves (position = index in `active`)
    std::vector<int> active;      // slot of each position
    int ncell_blk = 0, nnode_blk = 0, nitems = 0;
    int* cell_solve = nullptr;
    int* node_solve = nullptr;
    int2* items_f = nullptr;
    int2* items_b = nullptr;
    size_t maps_cap = 0;
    // per-iteration solve tables
    void* table_host = nullptr;   // pinned
    void* table_dev = nullptr;
    size_t table_cap = 0;         // bytes
    hipEvent_t table_copied = nullptr;
    bool table_pending = false;
    int sweep_grid = 0;
    int compute_units = 0;
    // ticket counters of the batch sweeps (forward, then backward), then the
    // block counters of the second stream's node values and gradient
    int* tickets = nullptr;
    // gradient_stream 2 (after the update) / 4 (after the next forward
    // sweep): the node values and gradient of a batch iteration on gstream;
    // gdone[k & 1] after iteration k's, which reads the state buffers that
    // the backward sweep of iteration k+2 overwrites
    int gradient_stream = 1;
    hipStream_t gstream = nullptr;
    hipEvent_t gstart = nullptr, gdone[2] = {nullptr, nullptr};
    bool gpend[2] = {false, false};
    long long iters = 0;          // batch iterations stepped (parity of gdone)
};

namespace {

constexpr int kBatchSweepWaves = 8;  // waves per sweep workgroup of a batch

template <typename R>
size_t batch_entry_bytes() {
    return sizeof(dev::BatchSolve<R>);
}

bool batch_ok(mfs_batch* b, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool batch_rebuild_maps(mfs_batch* b) {
    b->active.clear();
    for (size_t i = 0; i < b->slots.size(); ++i)
        if (b->slots[i]) b->active.push_back((int)i);
    std::vector<int> cs, ns;
    std::vector<int2> itf, itb;
    for (size_t a = 0; a < b->active.size(); ++a) {
        mfs_ctx* c = b->slots[b->active[a]];
        cs.insert(cs.end(), (size_t)c->nblk_cell, (int)a);
        ns.insert(ns.end(), (size_t)nblk_xcd(c->nN), (int)a);
    }
    // Flat schedule across solves; each solve's chunks stay in its sweep
    // order.  A solve's level order ends in a long thin tail (the cells on
    // its longest dependency chains, a few chunks per level).  The waves
    // that take tail chunks wait a level's latency each, so the tails of all
    // solves must run side by side: a tail meeting only other solves' wide
    // levels holds the waves in flight and the list behind it.  The tail is
    // last in the forward sweep and first in the backward sweep, so the
    // forward list aligns the solves at their last chunk (solve a starts at
    // step maxch - nchunk_a) and the backward list at their first.
    int maxch = 0;
    for (size_t a = 0; a < b->active.size(); ++a) maxch = std::max(maxch, b->slots[b->active[a]]->nchunk);
    for (int k = 0; k < maxch; ++k)
        for (size_t a = 0; a < b->active.size(); ++a) {
            const int nch = b->slots[b->active[a]]->nchunk;
            if (k < nch) itb.push_back(make_int2((int)a, k));
            if (k >= maxch - nch) itf.push_back(make_int2((int)a, k - (maxch - nch)));
        }
    b->ncell_blk = (int)cs.size();
    b->nnode_blk = (int)ns.size();
    b->nitems = (int)itf.size();
    if (!batch_ok(b, hipStreamSynchronize(b->stream), "batch sync")) return false;
    const size_t need = std::max({cs.size(), ns.size(), itf.size(), (size_t)1});
    if (need > b->maps_cap) {
        (void)hipFree(b->cell_solve);
        (void)hipFree(b->node_solve);
        (void)hipFree(b->items_f);
        (void)hipFree(b->items_b);
        b->cell_solve = b->node_solve = nullptr;
        b->items_f = b->items_b = nullptr;
        const size_t cap = need + need / 2;
        if (!batch_ok(b, hipMalloc((void**)&b->cell_solve, cap * sizeof(int)), "hipMalloc(batch maps)") ||
            !batch_ok(b, hipMalloc((void**)&b->node_solve, cap * sizeof(int)), "hipMalloc(batch maps)") ||
            !batch_ok(b, hipMalloc((void**)&b->items_f, cap * sizeof(int2)), "hipMalloc(batch maps)") ||
            !batch_ok(b, hipMalloc((void**)&b->items_b, cap * sizeof(int2)), "hipMalloc(batch maps)"))
            return false;
        b->maps_cap = cap;
    }
    if (!cs.empty() &&
        (!batch_ok(b, hipMemcpy(b->cell_solve, cs.data(), cs.size() * sizeof(int), hipMemcpyHostToDevice), "maps") ||
         !batch_ok(b, hipMemcpy(b->node_solve, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice), "maps") ||
         !batch_ok(b, hipMemcpy(b->items_f, itf.data(), itf.size() * sizeof(int2), hipMemcpyHostToDevice), "maps") ||
         !batch_ok(b, hipMemcpy(b->items_b, itb.data(), itb.size() * sizeof(int2), hipMemcpyHostToDevice), "maps")))
        return false;
    b->dirty = false;
    return true;
}

// The table entry of one solve for one iteration: `cur` is the state buffer
// the iteration starts from.
template <typename R>
void fill_entry(dev::BatchSolve<R>& E, mfs_ctx* c, int cur, int blk0_cell, int blk0_node, int iter_done, int slot) {
    std::memset(&E, 0, sizeof(E));
    const bool f64 = sizeof(R) == sizeof(double);
    E.nC = c->nC;
    E.nN = c->nN;
    E.n = c->nC + c->nB;
    E.blk0_cell = blk0_cell;
    E.nblk_cell = c->nblk_cell;
    E.blk0_node = blk0_node;
    E.nblk_node = (int)nblk_xcd(c->nN);
    E.cfo = c->cfo;
    E.cf = c->cf;
    E.fcell = c->fcell;
    E.fgeo = c->fgeo;
    E.fdn = c->fdn;
    E.vol = c->vol;
    E.q = c->q[cur];
    E.qnew = c->q[1 - cur];
    E.res = c->res;
    E.dt = c->dt;
    E.diag = c->diag;
    E.part = c->part;
    E.run = c->run;
    E.hist = c->hist;
    E.hist_slot = slot;
    E.cfl = cfl_of_iter(c->P, iter_done);
    E.cfl_min = 0.5 * E.cfl;
    E.ratio_dtmax = c->P.ratio_dtmax;
    E.q32 = c->q32;
    E.diag32 = c->diag32;
    E.K = c->K;
    E.snode = c->snode;
    E.sgeo64 = c->sgeo;
    E.sdn = c->sdn;
    E.sflip = c->sflip;
    E.nlo = c->nlo;
    E.ell_off = c->ell_off;
    E.ell_ab = c->ell_ab;
    E.nden = c->nden;
    E.fno = c->fno;
    E.fn = c->fn;
    E.btype = c->btype;
    auto& A = E.sw;
    A.nC = c->nC;
    A.nchunks = c->nchunk;
    A.nq = c->nC + c->nB;
    A.ntiles = c->ntiles;
    A.tile_chunk = c->tile_chunk;
    A.pos = c->pos;
    A.flat = 1;
    A.crun = c->crun;
    A.snb = c->snb;
    A.q64 = c->q[cur];
    A.res = c->res;
    A.qout = c->q[1 - cur];
    A.depmask = c->depmask;
    A.cell0 = c->cell0;
    A.err = c->err;
    A.tickets = A.tickets_next = nullptr;  // the batch sweeps assign (solve, position) items instead
    A.ngroups = 0;
    A.ubeg = 0;
    A.uend = c->nchunk;
    A.run_shift = 0;
    if constexpr (sizeof(R) == sizeof(double)) {
        E.ell_w = c->ell_w;
        E.qn = c->qn;
        E.fgeo_r = c->fgeo;
        E.gx = c->gx;
        E.gy = c->gy;
        E.gz = c->gz;
        A.sgeo = c->sgeo;
        A.q = c->q[cur];
        A.diag = c->diag;
        A.dqf = c->dqf;
        A.dqb = c->dq;
    } else {
        E.ell_w = c->ell_w32;
        E.qn = c->qn32;
        E.fgeo_r = c->fgeo32;
        E.gx = c->gx32;
        E.gy = c->gy32;
        E.gz = c->gz32;
        A.sgeo = c->sgeo32;
        A.q = c->q32;
        A.diag = c->diag32;
        A.dqf = c->dqf32;
        A.dqb = c->dq32;
    }
    (void)f64;
}

// Node values (with the symmetry projection) and the gradient of the
// batch iteration whose table is T, on stream st.
template <int MAXF, typename R>
bool batch_gradient(mfs_batch* b, const dev::BatchSolve<R>* T, hipStream_t st) {
    constexpr bool MIXED = sizeof(R) == sizeof(float);
    // on the second stream: one workgroup per CU (a multiple of 8) claiming blocks
    const int gw = ((b->compute_units + dev::kXcd - 1) / dev::kXcd) * dev::kXcd;
    int* const claim = b->tickets + kGradClaimOffset;
    const bool side = st != b->stream && gw > 0;
    if (side && gw < b->nnode_blk)
        hipLaunchKernelGGL((dev::k_b_node_claim<R>), dim3(gw), dim3(kBlock), 0, st, T, b->node_solve, b->nnode_blk,
                           claim);
    else
        hipLaunchKernelGGL((dev::k_b_node<R>), dim3(b->nnode_blk), dim3(kBlock), 0, st, T, b->node_solve);
    for (size_t a = 0; a < b->active.size(); ++a) {
        mfs_ctx* c = b->slots[b->active[a]];
        if (c->nsym > 0)
            hipLaunchKernelGGL(dev::k_node_symm<R>, dim3(nblk(c->nsym)), dim3(kBlock), 0, st, c->nsym, c->sym_node,
                               c->sym_n, MIXED ? (R*)c->qn32 : (R*)c->qn);
    }
    if (side && gw < b->ncell_blk)
        hipLaunchKernelGGL((dev::k_b_grad_claim<MAXF, R>), dim3(gw), dim3(kBlock), 0, st, T, b->cell_solve,
                           b->ncell_blk, claim + (kNodeClaimOffset - kGradClaimOffset));
    else
        hipLaunchKernelGGL((dev::k_b_grad<MAXF, R>), dim3(b->ncell_blk), dim3(kBlock), 0, st, T, b->cell_solve);
    return batch_ok(b, hipGetLastError(), "batch gradient");
}

template <int MAXF, int LPC, typename R>
int batch_step_t(mfs_batch* b, int nsteps) {
    constexpr bool MIXED = sizeof(R) == sizeof(float);
    const int ns = (int)b->active.size();
    if (ns == 0 || nsteps == 0) return MFS_OK;
    for (int a = 0; a < ns; ++a) {
        mfs_ctx* c = b->slots[b->active[a]];
        if (c->broken) {
            set_error("mfs_batch_step: a solve of the batch is in an error state");
            return MFS_ERR_STATE;
        }
        if (!ensure_hist(c, c->hist_n + nsteps)) return MFS_ERR_HIP;
    }
    // one table per iteration, read by every launch of that iteration
    const size_t per_iter = (size_t)ns * sizeof(dev::BatchSolve<R>);
    const size_t bytes = per_iter * (size_t)nsteps;
    if (b->table_pending) {
        if (!batch_ok(b, hipEventSynchronize(b->table_copied), "batch table event")) return MFS_ERR_HIP;
        b->table_pending = false;
    }
    if (bytes > b->table_cap) {
        if (!batch_ok(b, hipStreamSynchronize(b->stream), "batch sync")) return MFS_ERR_HIP;
        (void)hipHostFree(b->table_host);
        (void)hipFree(b->table_dev);
        b->table_host = b->table_dev = nullptr;
        b->table_cap = 0;
        const size_t cap = bytes + bytes / 2;
        if (!batch_ok(b, hipHostMalloc(&b->table_host, cap), "hipHostMalloc(batch table)") ||
            !batch_ok(b, hipMalloc(&b->table_dev, cap), "hipMalloc(batch table)"))
            return MFS_ERR_NOMEM;
        b->table_cap = cap;
    }
    auto* H = static_cast<dev::BatchSolve<R>*>(b->table_host);
    for (int i = 0; i < nsteps; ++i) {
        int blk_c = 0, blk_n = 0;
        for (int a = 0; a < ns; ++a) {
            mfs_ctx* c = b->slots[b->active[a]];
            const int cur = c->cur ^ (i & 1);
            fill_entry<R>(H[(size_t)i * ns + a], c, cur, blk_c, blk_n, c->iter_done + i, c->hist_n + i);
            blk_c += c->nblk_cell;
            blk_n += (int)nblk_xcd(c->nN);
        }
    }
    if (!batch_ok(b, hipMemcpyAsync(b->table_dev, b->table_host, bytes, hipMemcpyHostToDevice, b->stream),
                  "batch table upload") ||
        !batch_ok(b, hipEventRecord(b->table_copied, b->stream), "batch table event"))
        return MFS_ERR_HIP;
    b->table_pending = true;
    auto* D = static_cast<dev::BatchSolve<R>*>(b->table_dev);
    const dim3 sweep_grid((unsigned)std::max(1, std::min(b->sweep_grid, (b->nitems + kBatchSweepWaves - 1) /
                                                                         kBatchSweepWaves)));
    // The gradient feeds nothing in an iteration, so with a second stream it
    // runs beside the next iteration (see step_iterations).
    const int mode = b->gstream ? b->gradient_stream : 1;
    const bool side = mode >= 2;
    const dev::BatchSolve<R>* grad_T = nullptr;  // table of the gradient not yet launched
    int grad_par = 0;
    auto side_gradient = [&]() -> bool {
        if (!grad_T) return true;
        if (!batch_ok(b, hipEventRecord(b->gstart, b->stream), "record gradient start") ||
            !batch_ok(b, hipStreamWaitEvent(b->gstream, b->gstart, 0), "gradient waits") ||
            !batch_gradient<MAXF, R>(b, grad_T, b->gstream) ||
            !batch_ok(b, hipEventRecord(b->gdone[grad_par], b->gstream), "record gradient"))
            return false;
        b->gpend[grad_par] = true;
        grad_T = nullptr;
        return true;
    };
    for (int i = 0; i < nsteps; ++i) {
        const dev::BatchSolve<R>* Tf = D + (size_t)i * ns;
        // after the first iteration the previous gradient launch did this
        // pass when fused
        if (i == 0 || side)
            hipLaunchKernelGGL((dev::k_b_flux_ts<MAXF, MIXED, R>), dim3(b->ncell_blk), dim3(kBlock), 0, b->stream,
                               Tf, b->cell_solve);
        hipLaunchKernelGGL((dev::k_b_reduce<R>), dim3(ns), dim3(dev::kReduceBlock), 0, b->stream, Tf);
        hipLaunchKernelGGL((dev::k_b_clamp<MAXF, R>), dim3(b->ncell_blk), dim3(kBlock), 0, b->stream, Tf,
                           b->cell_solve);
        if (!batch_ok(b, hipGetLastError(), "batch flux/time step")) return MFS_ERR_HIP;
        const int par = (int)(b->iters & 1);
        if (side && b->gpend[par]) {
            if (!batch_ok(b, hipStreamWaitEvent(b->stream, b->gdone[par], 0), "wait gradient")) return MFS_ERR_HIP;
            b->gpend[par] = false;
        }
        {
            SweepTurn turn(b->device, b->stream);
            if (turn.error()) return sweep_turn_failed(nullptr, turn.error()) ? MFS_OK : MFS_ERR_HIP;
            int n_it = b->nitems;
            int ng = std::min(kTicketGroups, (int)sweep_grid.x);
            int* tk_f = b->tickets;
            int* tk_b = b->tickets + kTicketGroups * kTicketStride;
            void* args[] = {(void*)&Tf, (void*)&b->items_f, (void*)&n_it, (void*)&tk_f, (void*)&tk_b, (void*)&ng};
            if (!batch_ok(b,
                          hipLaunchKernel((const void*)dev::k_b_sweep<true, LPC, R>, sweep_grid,
                                                     dim3(64 * kBatchSweepWaves), args, 0, b->stream),
                          "batch forward sweep"))
                return MFS_ERR_HIP;
            if (side && mode == 4 && !side_gradient()) return MFS_ERR_HIP;
            void* argsb[] = {(void*)&Tf, (void*)&b->items_b, (void*)&n_it, (void*)&tk_b, (void*)&tk_f, (void*)&ng};
            if (!batch_ok(b,
                          hipLaunchKernel((const void*)dev::k_b_sweep<false, LPC, R>, sweep_grid,
                                                     dim3(64 * kBatchSweepWaves), argsb, 0, b->stream),
                          "batch backward sweep"))
                return MFS_ERR_HIP;
            if (!turn.done()) return sweep_turn_failed(nullptr, turn.error()) ? MFS_OK : MFS_ERR_HIP;
        }
        for (int a = 0; a < ns; ++a) {  // SetGhostVariables for the solves that asked for it
            mfs_ctx* c = b->slots[b->active[a]];
            if (c->ghost_refresh && c->nB > 0) {
                const int cur = c->cur ^ (i & 1);
                const mfs_params& P = c->P;
                const dev::GhostConst G{P.rho_s, P.u_s, P.v_s, P.w_s, P.p_s};
                hipLaunchKernelGGL(dev::k_ghost, dim3(nblk(c->nB)), dim3(kBlock), 0, b->stream, c->nB, c->nC,
                                   c->gface, c->fcell, c->fgeo, c->btype, c->K, G, c->q[1 - cur],
                                   MIXED ? c->q32 : nullptr);
            }
        }
        if (side) {
            grad_T = Tf;
            grad_par = par;
            if ((mode == 2 || i + 1 == nsteps) && !side_gradient()) return MFS_ERR_HIP;
        } else {
            hipLaunchKernelGGL((dev::k_b_node<R>), dim3(b->nnode_blk), dim3(kBlock), 0, b->stream, Tf,
                               b->node_solve);
            for (int a = 0; a < ns; ++a) {  // symmetry-plane projection, only where a solve has SYMM nodes
                mfs_ctx* c = b->slots[b->active[a]];
                if (c->nsym > 0)
                    hipLaunchKernelGGL(dev::k_node_symm<R>, dim3(nblk(c->nsym)), dim3(kBlock), 0, b->stream,
                                       c->nsym, c->sym_node, c->sym_n, MIXED ? (R*)c->qn32 : (R*)c->qn);
            }
            if (i + 1 < nsteps)  // the gradient with the next iteration's residual / time-step pass
                hipLaunchKernelGGL((dev::k_b_grad_flux_ts<MAXF, MIXED, R>), dim3(b->ncell_blk), dim3(kBlock), 0,
                                   b->stream, Tf, Tf + (size_t)ns, b->cell_solve);
            else
                hipLaunchKernelGGL((dev::k_b_grad<MAXF, R>), dim3(b->ncell_blk), dim3(kBlock), 0, b->stream, Tf,
                                   b->cell_solve);
            if (!batch_ok(b, hipGetLastError(), "batch gradient")) return MFS_ERR_HIP;
        }
        ++b->iters;
    }
    if (side) {  // join: later work on the batch stream comes after the last gradient
        const int last = (int)((b->iters - 1) & 1);
        if (!batch_ok(b, hipStreamWaitEvent(b->stream, b->gdone[last], 0), "join gradient")) return MFS_ERR_HIP;
        b->gpend[0] = b->gpend[1] = false;
    }
    for (int a = 0; a < ns; ++a) {
        mfs_ctx* c = b->slots[b->active[a]];
        const double cfl = cfl_of_iter(c->P, c->iter_done + nsteps - 1);
        c->cfl_min = 0.5 * cfl;
        c->cur ^= (nsteps & 1);
        c->hist_n += nsteps;
        c->iter_done += nsteps;
    }
    return MFS_OK;
}

template <int LPC, typename R>
int batch_sweep_occupancy(int* per_cu) {
    int a = 0, c = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, dev::k_b_sweep<true, LPC, R>, 64 * kBatchSweepWaves, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, dev::k_b_sweep<false, LPC, R>, 64 * kBatchSweepWaves, 0) !=
            hipSuccess)
        return -1;
    *per_cu = std::min(a, c);
    return 0;
}

}  // namespace

extern "C" int mfs_batch_create(int precision, mfs_batch** out) { return mfs_batch_create_ex(precision, nullptr, out); }

extern "C" int mfs_batch_create_ex(int precision, const mfs_batch_options* opt, mfs_batch** out) {
    if (!out || (precision != MFS_FP64 && precision != MFS_MIXED)) {
        set_error("mfs_batch_create: null argument or unknown precision");
        return MFS_ERR_ARG;
    }
    const int gradient_stream = (opt && opt->gradient_stream) ? opt->gradient_stream
                                : precision == MFS_MIXED   ? dev::kDefaultBatchGradientStreamMixed
                                                           : dev::kDefaultBatchGradientStreamFP64;
    if (gradient_stream != 1 && gradient_stream != 2 && gradient_stream != 4) {
        set_error("mfs_batch_create: gradient_stream must be 0 (default), 1, 2 or 4");
        return MFS_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error("mfs_batch_create: no HIP device available");
        return MFS_ERR_HIP;
    }
    mfs_batch* b = new mfs_batch();
    b->precision = precision;
    b->gradient_stream = gradient_stream;
    (void)hipGetDevice(&b->device);
    if (!sweep_order_setup(b->device)) {
        delete b;
        return MFS_ERR_HIP;
    }
    const size_t tbytes = sizeof(int) * kTicketInts;
    bool ok = batch_ok(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate") &&
              batch_ok(b, hipEventCreateWithFlags(&b->table_copied, hipEventDisableTiming), "hipEventCreate") &&
              batch_ok(b, hipMalloc(&b->tickets, tbytes), "hipMalloc(batch tickets)") &&
              batch_ok(b, hipMemsetAsync(b->tickets, 0, tbytes, b->stream), "memset batch tickets");
    if (ok && gradient_stream >= 2)
        ok = batch_ok(b, hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking), "hipStreamCreate(gradient)") &&
             batch_ok(b, hipEventCreateWithFlags(&b->gstart, hipEventDisableTiming), "hipEventCreate") &&
             batch_ok(b, hipEventCreateWithFlags(&b->gdone[0], hipEventDisableTiming), "hipEventCreate") &&
             batch_ok(b, hipEventCreateWithFlags(&b->gdone[1], hipEventDisableTiming), "hipEventCreate");
    if (!ok) {
        if (b->tickets) (void)hipFree(b->tickets);
        for (hipEvent_t e : {b->table_copied, b->gstart, b->gdone[0], b->gdone[1]})
            if (e) (void)hipEventDestroy(e);
        if (b->gstream) (void)hipStreamDestroy(b->gstream);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return MFS_ERR_HIP;
    }
    *out = b;
    return MFS_OK;
}

extern "C" int mfs_batch_attach(mfs_batch* b, mfs_ctx* c, int* slot) {
    if (!b || !c || !slot) {
        set_error("mfs_batch_attach: null argument");
        return MFS_ERR_ARG;
    }
    if (!c->owns_stream) {
        set_error("mfs_batch_attach: the context already belongs to a batch");
        return MFS_ERR_ARG;
    }
    if (c->precision != b->precision || c->device != b->device) {
        set_error("mfs_batch_attach: the context's precision or device differs from the batch's");
        return MFS_ERR_ARG;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        set_error("mfs_batch_attach: hipSetDevice failed");
        return MFS_ERR_HIP;
    }
    const bool first = b->maxf == 0;
    if (!first && (c->maxf != b->maxf || c->lpc != b->lpc)) {
        set_error("mfs_batch_attach: a batch holds either tetrahedral meshes or meshes with up to 6 faces per cell, not both");
        return MFS_ERR_MESH;
    }
    if (first) {
        int per_cu = 0, cus = 0;
        const bool mixed = b->precision == MFS_MIXED;
        const int occ = c->lpc == 4 ? (mixed ? batch_sweep_occupancy<4, float>(&per_cu)
                                             : batch_sweep_occupancy<4, double>(&per_cu))
                                    : (mixed ? batch_sweep_occupancy<8, float>(&per_cu)
                                             : batch_sweep_occupancy<8, double>(&per_cu));
        if (occ != 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) {
            set_error("mfs_batch_attach: occupancy query failed");
            return MFS_ERR_HIP;
        }
        b->maxf = c->maxf;
        b->lpc = c->lpc;
        b->sweep_grid = std::max(1, std::min(per_cu, kSweepWavesPerCU / kBatchSweepWaves) * cus);
        b->compute_units = cus;
    }
    // the solve's work goes to the batch stream from now on
    if (!batch_ok(b, hipStreamSynchronize(c->stream), "mfs_batch_attach: context stream")) return MFS_ERR_HIP;
    forget_sweep_stream(c->device, c->stream);
    (void)hipStreamDestroy(c->stream);
    c->stream = b->stream;
    c->owns_stream = false;
    int s = -1;
    for (size_t i = 0; i < b->slots.size(); ++i)
        if (!b->slots[i]) {
            s = (int)i;
            break;
        }
    if (s < 0) {
        s = (int)b->slots.size();
        b->slots.push_back(nullptr);
    }
    b->slots[s] = c;
    b->dirty = true;
    *slot = s;
    return MFS_OK;
}

extern "C" int mfs_batch_add(mfs_batch* b, const mfs_mesh_view* mesh, const mfs_params* p, int* slot) {
    if (!b || !mesh || !p || !slot) {
        set_error("mfs_batch_add: null argument");
        return MFS_ERR_ARG;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        set_error("mfs_batch_add: hipSetDevice failed");
        return MFS_ERR_HIP;
    }
    mfs_ctx* c = nullptr;
    int r = mfs_ctx_create(mesh, p, b->precision, &c);
    if (r != MFS_OK) return r;
    r = mfs_batch_attach(b, c, slot);
    if (r != MFS_OK) (void)mfs_ctx_destroy(c);
    return r;
}

extern "C" int mfs_batch_remove(mfs_batch* b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot]) {
        set_error("mfs_batch_remove: no such solve");
        return MFS_ERR_ARG;
    }
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    int r = mfs_ctx_destroy(b->slots[slot]);
    b->slots[slot] = nullptr;
    b->dirty = true;
    return r;
}

extern "C" int mfs_batch_step(mfs_batch* b, int nsteps) {
    if (!b || nsteps < 0) {
        set_error("mfs_batch_step: null batch or negative step count");
        return MFS_ERR_ARG;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        set_error("mfs_batch_step: hipSetDevice failed");
        return MFS_ERR_HIP;
    }
    if (b->dirty && !batch_rebuild_maps(b)) return MFS_ERR_HIP;
    if (b->precision == MFS_MIXED)
        return b->lpc == 4 ? batch_step_t<4, 4, float>(b, nsteps) : batch_step_t<6, 8, float>(b, nsteps);
    return b->lpc == 4 ? batch_step_t<4, 4, double>(b, nsteps) : batch_step_t<6, 8, double>(b, nsteps);
}

extern "C" int mfs_batch_sync(mfs_batch* b) {
    if (!b) return MFS_ERR_ARG;
    (void)hipSetDevice(b->device);
    if (!batch_ok(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize")) return MFS_ERR_HIP;
    for (mfs_ctx* c : b->slots) {
        if (!c) continue;
        int r = mfs_sync(c);
        if (r != MFS_OK) return r;
    }
    return MFS_OK;
}

extern "C" mfs_ctx* mfs_batch_ctx(mfs_batch* b, int slot) {
    if (!b || slot < 0 || slot >= (int)b->slots.size()) return nullptr;
    return b->slots[slot];
}

extern "C" int mfs_batch_count(mfs_batch* b) {
    if (!b) return 0;
    int n = 0;
    for (mfs_ctx* c : b->slots) n += c ? 1 : 0;
    return n;
}

extern "C" int mfs_batch_destroy(mfs_batch* b) {
    if (!b) return MFS_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    for (mfs_ctx*& c : b->slots)
        if (c) {
            (void)mfs_ctx_destroy(c);
            c = nullptr;
        }
    (void)hipFree(b->cell_solve);
    (void)hipFree(b->node_solve);
    (void)hipFree(b->items_f);
    (void)hipFree(b->items_b);
    (void)hipFree(b->tickets);
    (void)hipHostFree(b->table_host);
    (void)hipFree(b->table_dev);
    for (hipEvent_t e : {b->table_copied, b->gstart, b->gdone[0], b->gdone[1]})
        if (e) (void)hipEventDestroy(e);
    if (b->gstream) (void)hipStreamDestroy(b->gstream);
    forget_sweep_stream(b->device, b->stream);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return MFS_OK;
}
