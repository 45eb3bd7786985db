// HIP kernels and device context of the MI355X solver path (gfx950).
//
// Layout: structure-of-arrays in the mesh's (TDTree) numbering.  State q is
// 5 blocks of nC+nB doubles (rho, u, v, w, gauge p; ghost cell of boundary
// face k at nC+k), double-buffered so the backward sweep can write the
// updated state while its neighbours still read the old one.  Faces carry
// (c1,c2), (nx,ny,nz,area) and the two cell-to-face distances; cells carry
// their faces in ascending face id.
//
// Every reference loop that scatters face contributions into cells is run
// here as a per-cell gather over the cell's faces in ascending face id,
// which adds the same terms in the same order as the reference's serial
// face loop: results are bitwise equal, with no atomics and no colouring.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "mfs_ctx.hpp"
#include "mfs_mesh.hpp"
#include "mfs_prep.hpp"

namespace mfs {
void set_error(const std::string& msg);  // mfs_capi.cpp
}

namespace mfs {
namespace dev {

constexpr int kBlock = 256;
constexpr int kSweepWave = 64;
constexpr unsigned kSpinLimit = 1u << 22;  // bound on dependency polls of one sweep wave

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Deterministic block sum of NV doubles per thread (fixed shuffle tree,
// then a fixed LDS combination of the block's waves).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
    __shared__ double s[kBlock / 64][NV];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += __shfl_down(v[j], off, 64);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) s[w][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double a = s[0][j];
            for (int k = 1; k < (int)(blockDim.x >> 6); ++k) a += s[k][j];
            out[j] = a;
        }
    }
}

__device__ __forceinline__ void block_minmax(double vmin, double vmax, double* out_min, double* out_max) {
    __shared__ double smin_[kBlock / 64], smax_[kBlock / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        vmin = smin(vmin, __shfl_down(vmin, off, 64));
        vmax = smax(vmax, __shfl_down(vmax, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        smin_[w] = vmin;
        smax_[w] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = smin_[0], b = smax_[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            a = smin(a, smin_[k]);
            b = smax(b, smax_[k]);
        }
        *out_min = a;
        *out_max = b;
    }
}

// ---------------------------------------------------------------------------
// Residual: InviscidFlux + RoeFlux_noprec + LoadFlux (cal_flux.cpp:84-192,
// 682-833, 1219-1254).  res[c] starts at 0 (after ZeroResiduals) or at its
// current value (ACCUM) and receives -F for faces where c is c1 and +F where
// c is c2, in ascending face order.  The block partial sums of res^2 feed
// the residual norm of main.cpp:294-317.
template <bool ACCUM>
__global__ __launch_bounds__(kBlock) void k_flux(int nC, int n, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                 const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                                 const double* __restrict__ q, double* __restrict__ res,
                                                 double* __restrict__ part, FlowConst K) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < nC) {
        if (ACCUM)
#pragma unroll
            for (int j = 0; j < 5; ++j) r[j] = res[(size_t)j * nC + c];
        const int k0 = cfo[c], k1 = cfo[c + 1];
        for (int k = k0; k < k1; ++k) {
            const int f = cf[k];
            const int2 cc = fcell[f];
            const double4 g = fgeo[f];
            double ql[5], qr[5], F[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                ql[j] = q[(size_t)j * n + cc.x];
                qr[j] = q[(size_t)j * n + cc.y];
            }
            roe_flux<double>(ql, qr, g.x, g.y, g.z, g.w, K.gamm1, K.p_bar, K.alf_l, kTINYd, F);
            if (cc.x == c) {
#pragma unroll
                for (int j = 0; j < 5; ++j) r[j] -= F[j];
            } else {
#pragma unroll
                for (int j = 0; j < 5; ++j) r[j] += F[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) res[(size_t)j * nC + c] = r[j];
    }
    double s[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = r[j] * r[j];
    block_sum<5>(s, part + (size_t)blockIdx.x * 8);
}

__global__ void k_fill(double* p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Local time step: TimeStepNormal_new + CellIsMG + LimitTimeStep
// (compute_timestep.cpp:758-823, 1315-1338, 836-918), fused with the
// LU-SGS diagonal Diag = vol/dt + sum 0.5*A*(|u.n|+a)*lhs_omga
// (forward_lusgs.cpp:94-99, 554-595) when WITH_DIAG.  Both use the cell's
// own |u.n| + a on every face, evaluated once per (cell, face).
template <bool WITH_DIAG>
__global__ __launch_bounds__(kBlock) void k_timestep(int nC, int n, const int* __restrict__ cfo,
                                                     const int* __restrict__ cf, const int2* __restrict__ fcell,
                                                     const double4* __restrict__ fgeo, const double2* __restrict__ fdn,
                                                     const double* __restrict__ q, const double* __restrict__ vol,
                                                     double* __restrict__ dt, double* __restrict__ diag,
                                                     double* __restrict__ part, FlowConst K, double cfl,
                                                     double cfl_min) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double vmin = kBIGd, vmax = 0.0;
    if (c < nC) {
        const double rho = q[c], u = q[(size_t)n + c], v = q[2 * (size_t)n + c], w = q[3 * (size_t)n + c],
                     p = q[4 * (size_t)n + c];
        const double c2tmp = K.gam * (p + K.p_bar) / rho;
        const double a = sqrt(c2tmp);
        double dtc = kBIGd;
        int det = 1;
        const int k0 = cfo[c], k1 = cfo[c + 1];
        for (int k = k0; k < k1; ++k) {
            const int f = cf[k];
            const int2 cc = fcell[f];
            const double4 g = fgeo[f];
            const double2 d = fdn[f];
            const double dn = (cc.x == c) ? d.x : d.y;
            double vn = u * g.x + v * g.y + w * g.z;
            vn = fabs(vn);
            const double eigv = vn + a;
            dtc = smin(dtc, dn / eigv);
            const double p1 = q[4 * (size_t)n + cc.x], p2 = q[4 * (size_t)n + cc.y];
            const double dp = fabs(p2 - p1) / (p2 + p1 + K.p_bar + K.p_bar);
            if (dp > 0.1) det = 0;
        }
        double cfl_tmp;
        if (p > K.p_break)
            cfl_tmp = cfl;
        else if (p < K.p_min)
            cfl_tmp = cfl_min;
        else
            cfl_tmp = (p - K.p_min) / (K.p_break - K.p_min) * (cfl - cfl_min) + cfl_min;
        if (!det) cfl_tmp = cfl_min;
        dtc *= cfl_tmp;
        dt[c] = dtc;
        vmin = dtc;
        vmax = dtc;
        if (WITH_DIAG) {
            double dg = vol[c] / dtc;
            for (int k = k0; k < k1; ++k) {
                const double4 g = fgeo[cf[k]];
                double vn = u * g.x + v * g.y + w * g.z;
                vn = fabs(vn);
                const double eig = vn + a;
                dg += 0.5 * g.w * eig * K.lhs_omga;
            }
            diag[c] = dg;
        }
    }
    block_minmax(vmin, vmax, part + (size_t)blockIdx.x * 8 + 5, part + (size_t)blockIdx.x * 8 + 6);
}

// Diag alone (forward_lusgs.cpp:94-99 + CalDiagLUSGS :554-595), from dt.
// Only cells whose dt exceeds *lim are recomputed when lim is given (the
// ratio_dtmax clamp of compute_timestep.cpp:929-945).
__global__ __launch_bounds__(kBlock) void k_diag(int nC, int n, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                 const double4* __restrict__ fgeo, const double* __restrict__ q,
                                                 const double* __restrict__ vol, double* __restrict__ dt,
                                                 double* __restrict__ diag, const double* __restrict__ run,
                                                 double ratio_dtmax, FlowConst K) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nC) return;
    double dtc = dt[c];
    if (run) {
        if (run[2] == 0.0) return;
        const double lim = ratio_dtmax * run[1];
        if (!(dtc > lim)) return;
        dtc = lim;
        dt[c] = dtc;
    }
    const double rho = q[c], u = q[(size_t)n + c], v = q[2 * (size_t)n + c], w = q[3 * (size_t)n + c],
                 p = q[4 * (size_t)n + c];
    const double a = sqrt(K.gam * (p + K.p_bar) / rho);
    double dg = vol[c] / dtc;
    for (int k = cfo[c]; k < cfo[c + 1]; ++k) {
        const double4 g = fgeo[cf[k]];
        double vn = u * g.x + v * g.y + w * g.z;
        vn = fabs(vn);
        dg += 0.5 * g.w * (vn + a) * K.lhs_omga;
    }
    diag[c] = dg;
}

// ---------------------------------------------------------------------------
// Reductions.  mode bits: 1 = residual norms -> run[3..7], 2 = dt extrema
// (running max/min of compute_timestep.cpp:914-918, clamp flag), 4 = write
// history record `slot` from run[], 8 = reset the sweep tickets.
__global__ __launch_bounds__(kBlock) void k_reduce(int nblk, const double* __restrict__ part, double* __restrict__ run,
                                                   double* __restrict__ hist, int slot, int mode, double ratio_dtmax,
                                                   unsigned int* tickets) {
    if (mode & 1) {
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int b = threadIdx.x; b < nblk; b += kBlock)
#pragma unroll
            for (int j = 0; j < 5; ++j) s[j] += part[(size_t)b * 8 + j];
        __shared__ double tot[5];
        block_sum<5>(s, tot);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int j = 0; j < 5; ++j) run[3 + j] = sqrt(tot[j]);
    }
    if (mode & 2) {
        double a = kBIGd, b = 0.0;
        for (int k = threadIdx.x; k < nblk; k += kBlock) {
            a = smin(a, part[(size_t)k * 8 + 5]);
            b = smax(b, part[(size_t)k * 8 + 6]);
        }
        __shared__ double mn, mx;
        block_minmax(a, b, &mn, &mx);
        __syncthreads();
        if (threadIdx.x == 0) {
            const double dmax = smax(run[0], mx);
            const double dmin = smin(run[1], mn);
            run[0] = dmax;
            run[1] = dmin;
            run[2] = (dmax / dmin > ratio_dtmax) ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (mode & 4) {
            double* h = hist + (size_t)slot * 7;
            for (int j = 0; j < 5; ++j) h[j] = run[3 + j];
            h[5] = run[0];
            h[6] = run[1];
        }
        if (mode & 8) {
            tickets[0] = 0u;
            tickets[1] = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// LU-SGS sweeps (SolveLUSGS3D, forward_lusgs.cpp:803-1249) with
// LUSGSLayer[i] = i: the forward sweep of cell c uses the final DQ of its
// lower-numbered face neighbours, the backward sweep that of its
// higher-numbered neighbours and of its ghost cells (DQ = 0).  Cells are
// visited in level order (sched); each 64-cell chunk is claimed through a
// ticket counter, so a wave only waits for chunks claimed (and therefore
// running) before it, and it publishes each finished cell with a release
// fence and a per-cell epoch flag.  The arithmetic per cell is the serial
// sweep's, in the serial face order, so DQ is bitwise equal to it.
__device__ __forceinline__ int flag_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag_store(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct SweepArgs {
    int nC, nB, n;
    const int* cfo;
    const int* cf;
    const int2* fcell;
    const double4* fgeo;
    const double* q;      // state read by the sweep
    const double* res;    // forward: initial DQ
    const double* diag;
    double* dq;
    double* qout;         // backward: updated state
    const int* sched;
    int* flags;
    unsigned int* ticket;
    int* err;
    int epoch;
};

template <bool FWD>
__global__ __launch_bounds__(kSweepWave) void k_sweep(SweepArgs A, FlowConst K) {
    unsigned t = 0;
    if (threadIdx.x == 0) t = atomicAdd(A.ticket, 1u);
    t = __shfl(t, 0, 64);
    const int pos = (int)t * kSweepWave + (int)threadIdx.x;
    const bool active = pos < A.nC;
    const int c = active ? A.sched[pos] : 0;
    const int n = A.n, nC = A.nC;

    int k0 = 0, k1 = 0;
    unsigned pend = 0;
    double qc[5], D[5], dg = 1.0;
    if (active) {
        k0 = A.cfo[c];
        k1 = A.cfo[c + 1];
        for (int k = k0; k < k1 && k - k0 < 32; ++k) {
            const int f = A.cf[k];
            if (f < A.nB) continue;
            const int2 cc = A.fcell[f];
            const int nb = (cc.x == c) ? cc.y : cc.x;
            if (FWD ? (nb < c) : (nb > c)) pend |= 1u << (k - k0);
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) qc[j] = A.q[(size_t)j * n + c];
        dg = A.diag[c];
        if (FWD) {
#pragma unroll
            for (int j = 0; j < 5; ++j) D[j] = A.res[(size_t)j * nC + c];
        }
    }
    bool done = !active;
    unsigned spins = 0;
    const int want = A.epoch;
    while (true) {
        if (!done && pend) {
            unsigned p = pend;
            while (p) {
                const int b = __builtin_ctz(p);
                p &= p - 1;
                const int f = A.cf[k0 + b];
                const int2 cc = A.fcell[f];
                const int nb = (cc.x == c) ? cc.y : cc.x;
                if (flag_load(A.flags + nb) >= want) pend &= ~(1u << b);
            }
        }
        const bool ready = !done && pend == 0;
        if (__ballot(ready)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (ready) {
                if (!FWD) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) D[j] = A.dq[(size_t)j * n + c];
                }
                double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                const bool skip_faces = FWD && (c == 0);  // forward_lusgs.cpp:803-804
                if (!skip_faces) {
                    for (int k = k0; k < k1; ++k) {
                        const int f = A.cf[k];
                        const int2 cc = A.fcell[f];
                        const int nb = (cc.x == c) ? cc.y : cc.x;
                        if (FWD ? !(nb < c) : !(nb > c)) continue;
                        const double4 g = A.fgeo[f];
                        double nx = g.x, ny = g.y, nz = g.z;
                        if (cc.y == c) {
                            nx = -nx;
                            ny = -ny;
                            nz = -nz;
                        }
                        double qn[5], dqn[5], fl[5];
#pragma unroll
                        for (int j = 0; j < 5; ++j) qn[j] = A.q[(size_t)j * n + nb];
                        if (nb < nC) {
#pragma unroll
                            for (int j = 0; j < 5; ++j) dqn[j] = A.dq[(size_t)j * n + nb];
                        } else {
#pragma unroll
                            for (int j = 0; j < 5; ++j) dqn[j] = 0.0;
                        }
                        flux_lusgs<double>(qn, dqn, nx, ny, nz, K.gam, K.p_bar, K.lhs_omga, fl);
                        if (FWD) {
                            const double tmp = 0.5 * g.w;
#pragma unroll
                            for (int j = 0; j < 5; ++j) D[j] -= tmp * fl[j];
                        } else {
                            const double tmp = g.w;
#pragma unroll
                            for (int j = 0; j < 5; ++j) acc[j] += tmp * fl[j];
                        }
                    }
                }
                if (FWD) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) D[j] /= dg;
                    if (c != 0) dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
                } else {
                    const double tmp = 2.0 * dg;
#pragma unroll
                    for (int j = 0; j < 5; ++j) D[j] -= acc[j] / tmp;
                    dq_limit<double>(qc, D, K.gamm1, K.rho_min, K.rho_max, K.p_min, K.p_max);
                }
#pragma unroll
                for (int j = 0; j < 5; ++j) A.dq[(size_t)j * n + c] = D[j];
                if (!FWD) {
                    double qnew[5];
#pragma unroll
                    for (int j = 0; j < 5; ++j) qnew[j] = qc[j];
                    update_cell(qnew, D, K);
#pragma unroll
                    for (int j = 0; j < 5; ++j) A.qout[(size_t)j * n + c] = qnew[j];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (ready) {
                flag_store(A.flags + c, want);
                done = true;
            }
        }
        if (!__ballot(!done)) break;
        if (++spins > kSpinLimit) {
            if (lane_id() == 0) atomicExch(A.err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// ---------------------------------------------------------------------------
// Node values (CompNodeVar3D_dist, cal_gradient.cpp:1922-2372): each node
// sums its weighted contributions in the reference's order, then divides
// by WeightNode + TINY.
__global__ __launch_bounds__(kBlock) void k_node(int nN, int n, const int* __restrict__ nlo,
                                                 const int2* __restrict__ nlab, const double* __restrict__ nlw,
                                                 const double* __restrict__ nden, const double* __restrict__ q,
                                                 double* __restrict__ qn) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= nN) return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = nlo[p]; k < nlo[p + 1]; ++k) {
        const int2 ab = nlab[k];
        const double w = nlw[k];
        if (ab.y < 0) {
#pragma unroll
            for (int j = 0; j < 5; ++j) acc[j] += q[(size_t)j * n + ab.x] * w;
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                double fq = q[(size_t)j * n + ab.x] + q[(size_t)j * n + ab.y];
                fq *= 0.5;
                acc[j] += fq * w;
            }
        }
    }
    const double d = nden[p];
#pragma unroll
    for (int j = 0; j < 5; ++j) qn[(size_t)j * nN + p] = acc[j] / d;
}

// Symmetry-plane velocity projection of the node values (cal_gradient.cpp:2332-2360).
__global__ void k_node_symm(int nsym, int nN, const int* __restrict__ sym_node, const double* __restrict__ sym_n,
                            double* __restrict__ qn) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsym) return;
    const int p = sym_node[s];
    const double xs = sym_n[3 * s], ys = sym_n[3 * s + 1], zs = sym_n[3 * s + 2];
    const double un = qn[(size_t)nN + p], vn = qn[2 * (size_t)nN + p], wn = qn[3 * (size_t)nN + p];
    const double vv = un * xs + vn * ys + wn * zs;
    qn[(size_t)nN + p] = un - vv * xs;
    qn[2 * (size_t)nN + p] = vn - vv * ys;
    qn[3 * (size_t)nN + p] = wn - vv * zs;
}

// Green-Gauss gradient (CompGradientQ_Gauss_Node, cal_gradient.cpp:1740-1795,
// 1889-1897): face values from the node values (interior faces, SYMM and
// INTERFACE boundary faces) or the face-centre average of cell and ghost.
__global__ __launch_bounds__(kBlock) void k_grad(int nC, int nB, int n, int nN, const int* __restrict__ cfo,
                                                 const int* __restrict__ cf, const int2* __restrict__ fcell,
                                                 const double4* __restrict__ fgeo, const int* __restrict__ fno,
                                                 const int* __restrict__ fn, const int* __restrict__ btype,
                                                 const double* __restrict__ q, const double* __restrict__ qn,
                                                 const double* __restrict__ vol, double* __restrict__ gx,
                                                 double* __restrict__ gy, double* __restrict__ gz) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nC) return;
    double ax[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ay[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, az[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = cfo[c]; k < cfo[c + 1]; ++k) {
        const int f = cf[k];
        const int2 cc = fcell[f];
        const double4 g = fgeo[f];
        double qs[5];
        bool node_mean = true;
        if (f < nB) {
            const int t = btype[f];
            node_mean = (t == BC_INTERFACE || t == BC_SYMM);
        }
        if (node_mean) {
            const int j0 = fno[f], j1 = fno[f + 1];
            const double cnt = (double)(j1 - j0);
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                double s = 0.0;
                for (int m = j0; m < j1; ++m) s += qn[(size_t)j * nN + fn[m]];
                s /= cnt;
                qs[j] = s;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) qs[j] = 0.5 * (q[(size_t)j * n + cc.x] + q[(size_t)j * n + cc.y]);
        }
        const bool minus = (f >= nB) && (cc.y == c);
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const double s = qs[j] * g.w;
            const double tx = s * g.x, ty = s * g.y, tz = s * g.z;
            if (minus) {
                ax[j] -= tx;
                ay[j] -= ty;
                az[j] -= tz;
            } else {
                ax[j] += tx;
                ay[j] += ty;
                az[j] += tz;
            }
        }
    }
    const double vc = vol[c];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        gx[(size_t)j * n + c] = ax[j] / vc;
        gy[(size_t)j * n + c] = ay[j] / vc;
        gz[(size_t)j * n + c] = az[j] / vc;
    }
}

}  // namespace dev

// ===========================================================================
// Context management and launch sequences.
namespace {

using dev::kBlock;
using dev::kSweepWave;

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

// Event bracketing of one launch group when timing is enabled.
struct Timed {
    mfs_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(mfs_ctx* c, const char* nm) : ctx(c), name(nm) {
        if (!ctx->timing) return;
        a = take();
        b = take();
        if (a) hipEventRecord(a, ctx->stream);
    }
    ~Timed() {
        if (!ctx->timing || !a || !b) return;
        hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({name, {a, b}});
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
        ctx->event_pool.push_back(pe.second.first);
        ctx->event_pool.push_back(pe.second.second);
    }
    ctx->pending.clear();
}

bool check_launch(mfs_ctx* ctx, const char* what) { return hip_ok(ctx, hipGetLastError(), what); }

double cfl_of_iter(const mfs_params& P, int iter_done) {
    // LimitTimeStep (compute_timestep.cpp:868-871)
    if (iter_done > P.cfl_nstep) return P.cfl_end;
    const double ratio = P.cfl_end / P.cfl_start;
    return P.cfl_start * pow(ratio, (double)iter_done / P.cfl_nstep);
}

bool ensure_hist(mfs_ctx* ctx, int need) {
    if (need <= ctx->hist_cap) return true;
    int cap = std::max(need, std::max(64, 2 * ctx->hist_cap));
    double* nh = nullptr;
    void* v = nullptr;
    if (!hip_ok(ctx, hipMalloc(&v, sizeof(double) * 7 * (size_t)cap), "hipMalloc(history)")) return false;
    nh = static_cast<double*>(v);
    if (ctx->hist && ctx->hist_n > 0)
        if (!hip_ok(ctx, hipMemcpyAsync(nh, ctx->hist, sizeof(double) * 7 * ctx->hist_n, hipMemcpyDeviceToDevice, ctx->stream),
                    "copy history"))
            return false;
    if (ctx->hist) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->hist);
    }
    ctx->hist = nh;
    ctx->hist_cap = cap;
    return true;
}

// ---- launch helpers
bool launch_flux(mfs_ctx* ctx, bool accum) {
    Timed t(ctx, "flux");
    const double* q = ctx->q[ctx->cur];
    const int n = ctx->nC + ctx->nB;
    if (accum)
        hipLaunchKernelGGL(dev::k_flux<true>, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n, ctx->cfo,
                           ctx->cf, ctx->fcell, ctx->fgeo, q, ctx->res, ctx->part, ctx->K);
    else
        hipLaunchKernelGGL(dev::k_flux<false>, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n, ctx->cfo,
                           ctx->cf, ctx->fcell, ctx->fgeo, q, ctx->res, ctx->part, ctx->K);
    return check_launch(ctx, "k_flux");
}

bool launch_timestep(mfs_ctx* ctx, bool with_diag, double cfl, double cfl_min) {
    Timed t(ctx, "timestep");
    const double* q = ctx->q[ctx->cur];
    const int n = ctx->nC + ctx->nB;
    if (with_diag)
        hipLaunchKernelGGL(dev::k_timestep<true>, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, ctx->fdn, q, ctx->vol, ctx->dt, ctx->diag,
                           ctx->part, ctx->K, cfl, cfl_min);
    else
        hipLaunchKernelGGL(dev::k_timestep<false>, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, ctx->fdn, q, ctx->vol, ctx->dt, ctx->diag,
                           ctx->part, ctx->K, cfl, cfl_min);
    return check_launch(ctx, "k_timestep");
}

bool launch_reduce(mfs_ctx* ctx, int mode, int slot) {
    Timed t(ctx, "reduce");
    hipLaunchKernelGGL(dev::k_reduce, dim3(1), dim3(kBlock), 0, ctx->stream, (int)ctx->nblk_cell, ctx->part, ctx->run,
                       ctx->hist, slot, mode, ctx->P.ratio_dtmax, ctx->tickets);
    return check_launch(ctx, "k_reduce");
}

bool launch_diag(mfs_ctx* ctx, bool clamp_only) {
    Timed t(ctx, "diag");
    const int n = ctx->nC + ctx->nB;
    hipLaunchKernelGGL(dev::k_diag, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n, ctx->cfo, ctx->cf,
                       ctx->fgeo, ctx->q[ctx->cur], ctx->vol, ctx->dt, ctx->diag, clamp_only ? ctx->run : nullptr,
                       ctx->P.ratio_dtmax, ctx->K);
    return check_launch(ctx, "k_diag");
}

bool launch_sweeps(mfs_ctx* ctx) {
    const int n = ctx->nC + ctx->nB;
    const unsigned nchunk = nblk(ctx->nC, kSweepWave);
    dev::SweepArgs A;
    A.nC = ctx->nC;
    A.nB = ctx->nB;
    A.n = n;
    A.cfo = ctx->cfo;
    A.cf = ctx->cf;
    A.fcell = ctx->fcell;
    A.fgeo = ctx->fgeo;
    A.q = ctx->q[ctx->cur];
    A.res = ctx->res;
    A.diag = ctx->diag;
    A.dq = ctx->dq;
    A.qout = ctx->q[1 - ctx->cur];
    A.flags = ctx->flags;
    A.err = ctx->err;
    {
        Timed t(ctx, "fwd_sweep");
        A.sched = ctx->sched_f;
        A.ticket = ctx->tickets;
        A.epoch = ctx->epoch + 1;
        hipLaunchKernelGGL(dev::k_sweep<true>, dim3(nchunk), dim3(kSweepWave), 0, ctx->stream, A, ctx->K);
        if (!check_launch(ctx, "k_sweep<fwd>")) return false;
    }
    {
        Timed t(ctx, "bwd_sweep");
        A.sched = ctx->sched_b;
        A.ticket = ctx->tickets + 1;
        A.epoch = ctx->epoch + 2;
        hipLaunchKernelGGL(dev::k_sweep<false>, dim3(nchunk), dim3(kSweepWave), 0, ctx->stream, A, ctx->K);
        if (!check_launch(ctx, "k_sweep<bwd>")) return false;
    }
    ctx->epoch += 2;
    ctx->cur = 1 - ctx->cur;
    return true;
}

bool launch_gradient(mfs_ctx* ctx) {
    const int n = ctx->nC + ctx->nB;
    const double* q = ctx->q[ctx->cur];
    {
        Timed t(ctx, "node");
        hipLaunchKernelGGL(dev::k_node, dim3(nblk(ctx->nN)), dim3(kBlock), 0, ctx->stream, ctx->nN, n, ctx->nlo, ctx->nlab,
                           ctx->nlw, ctx->nden, q, ctx->qn);
        if (!check_launch(ctx, "k_node")) return false;
        if (ctx->nsym > 0) {
            hipLaunchKernelGGL(dev::k_node_symm, dim3(nblk(ctx->nsym)), dim3(kBlock), 0, ctx->stream, ctx->nsym, ctx->nN,
                               ctx->sym_node, ctx->sym_n, ctx->qn);
            if (!check_launch(ctx, "k_node_symm")) return false;
        }
    }
    {
        Timed t(ctx, "gradient");
        hipLaunchKernelGGL(dev::k_grad, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, ctx->nB, n, ctx->nN,
                           ctx->cfo, ctx->cf, ctx->fcell, ctx->fgeo, ctx->fno, ctx->fn, ctx->btype, q, ctx->qn, ctx->vol,
                           ctx->gx, ctx->gy, ctx->gz);
        if (!check_launch(ctx, "k_grad")) return false;
    }
    return true;
}

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
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs) hipFree(b.p);
    ctx->bufs.clear();
    if (ctx->hist) hipFree(ctx->hist);
    ctx->hist = nullptr;
    for (auto& pe : ctx->pending) {
        hipEventDestroy(pe.second.first);
        hipEventDestroy(pe.second.second);
    }
    ctx->pending.clear();
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    ctx->event_pool.clear();
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    ctx->stream = nullptr;
}

}  // namespace
}  // namespace mfs

// ===========================================================================
// C ABI: context functions.
using namespace mfs;

extern "C" int mfs_ctx_create(const mfs_mesh_view* mesh, const mfs_params* p, int precision, mfs_ctx** out) {
    if (!mesh || !p || !out) {
        set_error("mfs_ctx_create: null argument");
        return MFS_ERR_ARG;
    }
    *out = nullptr;
    if (precision != MFS_FP64) {
        set_error("mfs_ctx_create: this build supports MFS_FP64 contexts only");
        return MFS_ERR_ARG;
    }
    HostLayout L;
    std::string e = build_layout(*mesh, L);
    if (!e.empty()) {
        set_error("mfs_ctx_create: " + e);
        return MFS_ERR_MESH;
    }
    for (int c = 0; c < L.nC; ++c)
        if (L.cfo[c + 1] - L.cfo[c] > 32) {
            set_error("mfs_ctx_create: cells with more than 32 faces are not supported");
            return MFS_ERR_MESH;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error("mfs_ctx_create: no HIP device available");
        return MFS_ERR_HIP;
    }
    mfs_ctx* ctx = new mfs_ctx();
    hipGetDevice(&ctx->device);
    ctx->nC = L.nC;
    ctx->nF = L.nF;
    ctx->nB = L.nB;
    ctx->nN = L.nN;
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
    ctx->nlev_f = L.nlev_f;
    ctx->nlev_b = L.nlev_b;
    ctx->nblk_cell = (int)nblk(L.nC);
    const size_t n = (size_t)L.nC + L.nB;
    bool ok = hip_ok(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    ok = ok && dalloc(ctx, ctx->q[0], 5 * n, "q0") && dalloc(ctx, ctx->q[1], 5 * n, "q1");
    ok = ok && dalloc(ctx, ctx->res, 5 * (size_t)L.nC, "res") && dalloc(ctx, ctx->dt, L.nC, "dt") &&
         dalloc(ctx, ctx->diag, L.nC, "diag") && dalloc(ctx, ctx->dq, 5 * n, "dq");
    ok = ok && dalloc(ctx, ctx->gx, 5 * n, "dqdx") && dalloc(ctx, ctx->gy, 5 * n, "dqdy") &&
         dalloc(ctx, ctx->gz, 5 * n, "dqdz") && dalloc(ctx, ctx->qn, 5 * (size_t)L.nN, "qnode");
    ok = ok && dalloc(ctx, ctx->fcell, L.nF, "fcell") && dalloc(ctx, ctx->fgeo, L.nF, "fgeo") &&
         dalloc(ctx, ctx->fdn, L.nF, "fdn") && dalloc(ctx, ctx->fno, L.nF + 1, "fno") &&
         dalloc(ctx, ctx->fn, L.fn.size(), "fn") && dalloc(ctx, ctx->btype, L.nB, "btype");
    ok = ok && dalloc(ctx, ctx->cfo, L.nC + 1, "cfo") && dalloc(ctx, ctx->cf, L.cf.size(), "cf") &&
         dalloc(ctx, ctx->vol, L.nC, "vol");
    ok = ok && dalloc(ctx, ctx->nlo, L.nN + 1, "nlo") && dalloc(ctx, ctx->nlab, L.nla.size(), "nlab") &&
         dalloc(ctx, ctx->nlw, L.nlw.size(), "nlw") && dalloc(ctx, ctx->nden, L.nN, "nden");
    ctx->nsym = (int)L.sym_node.size();
    ok = ok && dalloc(ctx, ctx->sym_node, L.sym_node.size(), "sym_node") &&
         dalloc(ctx, ctx->sym_n, L.sym_n.size(), "sym_n");
    ok = ok && dalloc(ctx, ctx->sched_f, L.nC, "sched_f") && dalloc(ctx, ctx->sched_b, L.nC, "sched_b") &&
         dalloc(ctx, ctx->flags, L.nC, "flags") && dalloc(ctx, ctx->tickets, 2, "tickets");
    ok = ok && dalloc(ctx, ctx->part, 8 * (size_t)ctx->nblk_cell, "partials") && dalloc(ctx, ctx->run, 8, "run") &&
         dalloc(ctx, ctx->err, 4, "err");
    ok = ok && ensure_hist(ctx, 64);
    if (ok) {
        std::vector<int2> fc(L.nF);
        std::vector<double2> fd(L.nF);
        for (int i = 0; i < L.nF; ++i) {
            fc[i] = make_int2(L.c1[i], L.c2[i]);
            fd[i] = make_double2(L.dn1[i], L.dn2[i]);
        }
        std::vector<int2> nab(L.nla.size());
        for (size_t k = 0; k < L.nla.size(); ++k) nab[k] = make_int2(L.nla[k], L.nlb[k]);
        ok = upload(ctx, ctx->q[0], L.q.data(), 5 * n, "upload q") && upload(ctx, ctx->q[1], L.q.data(), 5 * n, "upload q") &&
             upload(ctx, ctx->fcell, fc.data(), L.nF, "upload fcell") &&
             upload(ctx, ctx->fgeo, reinterpret_cast<const double4*>(L.fgeo.data()), L.nF, "upload fgeo") &&
             upload(ctx, ctx->fdn, fd.data(), L.nF, "upload fdn") && upload(ctx, ctx->fno, L.fno.data(), L.nF + 1, "upload fno") &&
             upload(ctx, ctx->fn, L.fn.data(), L.fn.size(), "upload fn") &&
             upload(ctx, ctx->btype, L.btype.data(), L.nB, "upload btype") &&
             upload(ctx, ctx->cfo, L.cfo.data(), L.nC + 1, "upload cfo") && upload(ctx, ctx->cf, L.cf.data(), L.cf.size(), "upload cf") &&
             upload(ctx, ctx->vol, L.vol.data(), L.nC, "upload vol") && upload(ctx, ctx->nlo, L.nlo.data(), L.nN + 1, "upload nlo") &&
             upload(ctx, ctx->nlab, nab.data(), nab.size(), "upload nlab") &&
             upload(ctx, ctx->nlw, L.nlw.data(), L.nlw.size(), "upload nlw") &&
             upload(ctx, ctx->nden, L.nden.data(), L.nN, "upload nden") &&
             upload(ctx, ctx->sym_node, L.sym_node.data(), L.sym_node.size(), "upload sym") &&
             upload(ctx, ctx->sym_n, L.sym_n.data(), L.sym_n.size(), "upload sym_n") &&
             upload(ctx, ctx->sched_f, L.sched_f.data(), L.nC, "upload sched_f") &&
             upload(ctx, ctx->sched_b, L.sched_b.data(), L.nC, "upload sched_b");
        const double run0[8] = {0.0, kBIGd, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        ok = ok && upload(ctx, ctx->run, run0, 8, "upload run");
        ok = ok && hip_ok(ctx, hipMemsetAsync(ctx->flags, 0, sizeof(int) * L.nC, ctx->stream), "memset flags") &&
             hip_ok(ctx, hipMemsetAsync(ctx->tickets, 0, sizeof(unsigned) * 2, ctx->stream), "memset tickets") &&
             hip_ok(ctx, hipMemsetAsync(ctx->err, 0, sizeof(int) * 4, ctx->stream), "memset err") &&
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
    hipSetDevice(ctx->device);
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
    // ratio_dtmax clamp (compute_timestep.cpp:929-945): dt only
    Timed t(ctx, "clamp");
    const int n = ctx->nC + ctx->nB;
    hipLaunchKernelGGL(dev::k_diag, dim3(ctx->nblk_cell), dim3(kBlock), 0, ctx->stream, ctx->nC, n, ctx->cfo, ctx->cf,
                       ctx->fgeo, ctx->q[ctx->cur], ctx->vol, ctx->dt, ctx->diag, ctx->run, ctx->P.ratio_dtmax, ctx->K);
    return check_launch(ctx, "k_diag(clamp)") ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_forward_lusgs(mfs_ctx* ctx) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (!launch_diag(ctx, false)) return MFS_ERR_HIP;
    if (!launch_reduce(ctx, 8, -1)) return MFS_ERR_HIP;
    return launch_sweeps(ctx) ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_pass_interface_data(mfs_ctx* ctx) { return usable(ctx) ? MFS_OK : MFS_ERR_STATE; }

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

extern "C" int mfs_step(mfs_ctx* ctx, int nsteps) {
    if (!usable(ctx)) return MFS_ERR_STATE;
    if (nsteps < 0) {
        set_error("mfs_step: negative step count");
        return MFS_ERR_ARG;
    }
    if (!ensure_hist(ctx, ctx->hist_n + nsteps)) return MFS_ERR_HIP;
    for (int s = 0; s < nsteps; ++s) {
        const double cfl = cfl_of_iter(ctx->P, ctx->iter_done);
        const double cfl_min = 0.5 * cfl;
        ctx->cfl_min = cfl_min;
        // ZeroResiduals + InviscidFlux: the flux kernel writes res from zero.
        if (!launch_flux(ctx, false)) return MFS_ERR_HIP;
        // ComputeTimeStep with the LU-SGS diagonal fused in.
        if (!launch_timestep(ctx, true, cfl, cfl_min)) return MFS_ERR_HIP;
        if (!launch_reduce(ctx, 1 | 2 | 4 | 8, ctx->hist_n)) return MFS_ERR_HIP;
        if (!launch_diag(ctx, true)) return MFS_ERR_HIP;  // only acts when the ratio_dtmax clamp triggers
        // ForwardLUSGS: sweeps + update
        if (!launch_sweeps(ctx)) return MFS_ERR_HIP;
        // CalculateGradient
        if (!launch_gradient(ctx)) return MFS_ERR_HIP;
        ctx->hist_n++;
        ctx->iter_done++;
    }
    return MFS_OK;
}

extern "C" int mfs_sync(mfs_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return MFS_ERR_ARG;
    }
    hipSetDevice(ctx->device);
    if (!hip_ok(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) return MFS_ERR_HIP;
    int err[4] = {0, 0, 0, 0};
    if (!hip_ok(ctx, hipMemcpy(err, ctx->err, sizeof(err), hipMemcpyDeviceToHost), "read error flags")) return MFS_ERR_HIP;
    if (ctx->timing) collect_timing(ctx);
    if (err[0]) {
        ctx->broken = true;
        set_error("LU-SGS sweep exceeded its dependency wait limit");
        return MFS_ERR_STATE;
    }
    return ctx->broken ? MFS_ERR_STATE : MFS_OK;
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

static double* field_ptr(mfs_ctx* ctx, int field, size_t* count, int* state_block) {
    const size_t n = (size_t)ctx->nC + ctx->nB;
    *state_block = -1;
    switch (field) {
        case MFS_F_RHO:
        case MFS_F_U:
        case MFS_F_V:
        case MFS_F_W:
        case MFS_F_P:
            *count = n;
            *state_block = field;
            return ctx->q[ctx->cur] + (size_t)field * n;
        case MFS_F_RES: *count = 5 * (size_t)ctx->nC; return ctx->res;
        case MFS_F_DT: *count = ctx->nC; return ctx->dt;
        case MFS_F_DQ: *count = 5 * n; return ctx->dq;
        case MFS_F_DQDX: *count = 5 * n; return ctx->gx;
        case MFS_F_DQDY: *count = 5 * n; return ctx->gy;
        case MFS_F_DQDZ: *count = 5 * n; return ctx->gz;
        case MFS_F_DIAG: *count = ctx->nC; return ctx->diag;
        case MFS_F_QNODE: *count = 5 * (size_t)ctx->nN; return ctx->qn;
        default: *count = 0; return nullptr;
    }
}

extern "C" size_t mfs_field_count(mfs_ctx* ctx, int field) {
    if (!ctx) return 0;
    size_t cnt = 0;
    int sb;
    field_ptr(ctx, field, &cnt, &sb);
    return cnt;
}

extern "C" int mfs_download(mfs_ctx* ctx, int field, double* dst, size_t count) {
    if (!ctx || !dst) {
        set_error("mfs_download: null argument");
        return MFS_ERR_ARG;
    }
    size_t cnt = 0;
    int sb;
    double* src = field_ptr(ctx, field, &cnt, &sb);
    if (!src || count != cnt) {
        set_error("mfs_download: unknown field or wrong count");
        return MFS_ERR_ARG;
    }
    int r = mfs_sync(ctx);
    if (r != MFS_OK) return r;
    return hip_ok(ctx, hipMemcpy(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost), "download") ? MFS_OK : MFS_ERR_HIP;
}

extern "C" int mfs_upload(mfs_ctx* ctx, int field, const double* src, size_t count) {
    if (!usable(ctx) || !src) return MFS_ERR_ARG;
    size_t cnt = 0;
    int sb;
    double* dst = field_ptr(ctx, field, &cnt, &sb);
    if (!dst || count != cnt) {
        set_error("mfs_upload: unknown field or wrong count");
        return MFS_ERR_ARG;
    }
    if (!hip_ok(ctx, hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, ctx->stream), "upload"))
        return MFS_ERR_HIP;
    if (sb >= 0) {
        const size_t n = (size_t)ctx->nC + ctx->nB;
        double* other = ctx->q[1 - ctx->cur] + (size_t)sb * n;
        if (!hip_ok(ctx, hipMemcpyAsync(other, src, sizeof(double) * cnt, hipMemcpyHostToDevice, ctx->stream), "upload"))
            return MFS_ERR_HIP;
    }
    return hip_ok(ctx, hipStreamSynchronize(ctx->stream), "upload sync") ? MFS_OK : MFS_ERR_HIP;
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
    info->sweep_levels = ctx->nlev_f;
    info->device_bytes = ctx->device_bytes;
    return MFS_OK;
}

extern "C" int mfs_ctx_enable_timing(mfs_ctx* ctx, int enable) {
    if (!ctx) return MFS_ERR_ARG;
    ctx->timing = enable != 0;
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
