// Per-cell and per-node gather kernels of the solver path, included by
// mfs_device.hip inside namespace mfs::dev.
//
// Device layout: state q, residual and node values are stored cell-major
// (5 consecutive doubles per cell or node) so that a gather of one
// neighbour touches one or two cache lines; faces and nodes are stored in
// first-touch order of the cell traversal.  Each kernel visits a cell's
// faces in ascending mesh face id (the per-cell list order), which is the
// order in which the reference's face loops add into that cell.
//
// The MAXF template bounds the faces per cell (4 for tetrahedral meshes, 6
// otherwise).  Slots beyond a cell's face count are loaded from the cell's
// first face, so every load of a cell can be issued at once, and their
// contributions are skipped.

// Real types of the kernels: FP64 (the reference's arithmetic) and FP32
// (the mixed-precision operator).  kSent is a NaN payload that arithmetic
// never produces, used by the sweeps to mark unpublished values.
template <typename R>
struct Real;
template <>
struct Real<double> {
    using bits = unsigned long long;
    using vec4 = double4;
    static constexpr bits kSent = 0x7FF4D51F5EED0000ULL;
};
template <>
struct Real<float> {
    using bits = unsigned int;
    using vec4 = float4;
    static constexpr bits kSent = 0x7FA5D51Fu;
};

template <int NV, int NW = kBlock / 64>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out, int stride = 1) {
    __shared__ double s[NW][NV];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += __shfl_down(v[j], off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) s[w][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double a = s[0][j];
            for (int k = 1; k < (int)(blockDim.x >> 6); ++k) a += s[k][j];
            out[(size_t)j * stride] = a;
        }
    }
}

template <int NW = kBlock / 64>
__device__ __forceinline__ void block_minmax(double vmin, double vmax, double* out_min, double* out_max) {
    __shared__ double smin_[NW], smax_[NW];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        vmin = smin(vmin, __shfl_down(vmin, off, 64));
        vmax = smax(vmax, __shfl_down(vmax, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
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

// Logical index of this workgroup.  Workgroups are dispatched round-robin
// over the 8 XCDs, each with its own L2; mapping consecutive logical blocks
// to one XCD gives each L2 a contiguous range of cells, whose neighbours
// (and their faces and nodes) are mostly in the same range.  The grid is a
// multiple of kXcd blocks.
__device__ __forceinline__ int xcd_block() {
    const int b = (int)blockIdx.x;
    return (b % kXcd) * ((int)gridDim.x / kXcd) + b / kXcd;
}

__device__ __forceinline__ void load5(const double* __restrict__ p, double v[5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) v[j] = p[j];
}

__global__ void k_fill(double* p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i(int* p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// SetGhostVariables (src/Kernels/load_grid.cpp:953-1212), inviscid branch:
// the ghost state behind each boundary face from the state of the cell in
// front of it.  One thread per boundary face k; the ghost is cell nC+k and
// gface[k] the device id of its face.  WALL and SYMM mirror the velocity
// in the face plane; FAR_FIELD takes the Riemann-invariant state
// (supersonic in/outflow, subsonic in/outflow with the entropy of the
// upwind side) and extrapolates it to the ghost.  Expressions and their
// order follow the reference; the FAR_FIELD subsonic branch calls pow,
// whose GPU result may differ from the host's in the last bit.
struct GhostConst {
    double rho_s, u_s, v_s, w_s, p_s;  // freestream (p gauge)
};

__global__ __launch_bounds__(kBlock) void k_ghost(int nB, int nC, const int* __restrict__ gface,
                                                  const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                                  const int* __restrict__ btype, FlowConst K, GhostConst G,
                                                  double* __restrict__ q, float* __restrict__ q32) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= nB) return;
    const int type = btype[k];
    if (type == BC_INTERFACE) return;
    const int f = gface[k];
    const int c1 = fcell[f].x;
    const size_t o1 = (size_t)c1 * 5, o2 = (size_t)(nC + k) * 5;
    const double4 g = fgeo[f];
    const double xfn = g.x, yfn = g.y, zfn = g.z;
    const double gam = K.gam, p_bar = K.p_bar;
    const double r1 = q[o1], u1 = q[o1 + 1], v1 = q[o1 + 2], w1 = q[o1 + 3], p1 = q[o1 + 4];
    double rho2 = q[o2], u2 = q[o2 + 1], v2 = q[o2 + 2], w2 = q[o2 + 3], p2 = q[o2 + 4];
    if (type == BC_WALL || type == BC_SYMM) {
        p2 = p1;
        rho2 = r1;
        const double vn = 2. * (xfn * u1 + yfn * v1 + zfn * w1);
        u2 = u1 - vn * xfn;
        v2 = v1 - vn * yfn;
        w2 = w1 - vn * zfn;
    } else if (type == BC_FAR_FIELD) {
        const double um = G.u_s, vm = G.v_s, wm = G.w_s;
        const double up = u1, vp = v1, wp = w1;
        const double rhom = G.rho_s, pm = G.p_s + p_bar;
        const double rhop = r1, pp = p1 + p_bar;
        const double norm_of_uvw = sqrt(um * um + vm * vm + wm * wm);
        const double vnm = xfn * um + yfn * vm + zfn * wm;
        const double vnp = xfn * up + yfn * vp + zfn * wp;
        const double cm = sqrt(gam * pm / rhom);
        const double cp = sqrt(gam * pp / rhop);
        const double riemm = vnm - 2. * cm / (gam - 1.);
        const double riemp = vnp + 2. * cp / (gam - 1.);
        const double vnb = 0.5 * (riemp + riemm);
        const double cb = 0.25 * (riemp - riemm) * (gam - 1.);
        if (fabs(vnb / cb) > 1.) {
            if (vnb <= 0.0) {
                rho2 = rhom;
                u2 = um;
                v2 = vm;
                w2 = wm;
                p2 = pm;
            } else {
                rho2 = rhop;
                u2 = up;
                v2 = vp;
                w2 = wp;
                p2 = pp;
            }
        } else {
            const double rela_vnb = vnb / norm_of_uvw;
            if (rela_vnb <= -0.0 || rela_vnb > 0.0) {
                const bool in = rela_vnb <= -0.0;
                const double bu = in ? um : up, bv = in ? vm : vp, bw = in ? wm : wp, bvn = in ? vnm : vnp;
                const double bp = in ? pm : pp, brho = in ? rhom : rhop;
                const double vtx = bu - bvn * xfn, vty = bv - bvn * yfn, vtz = bw - bvn * zfn;
                const double entr = bp / pow(brho, gam);
                rho2 = pow((cb * cb / (entr * gam)), double(1. / (gam - 1.)));
                u2 = vtx + vnb * xfn;
                v2 = vty + vnb * yfn;
                w2 = vtz + vnb * zfn;
                p2 = cb * cb * rho2 / gam;
            } else {
                rho2 = 0.5 * (rhop + rhom);
                u2 = 0.5 * (up + um);
                v2 = 0.5 * (vp + vm);
                w2 = 0.5 * (wp + wm);
                p2 = 0.5 * (pp + pm);
            }
        }
        rho2 = 2 * rho2 - rhop;
        u2 = 2 * u2 - up;
        v2 = 2 * v2 - vp;
        w2 = 2 * w2 - wp;
        p2 = 2 * p2 - pp;
        p2 -= p_bar;
    }  // other types: the reference reports an error, keeps the ghost state and clamps it below
    rho2 = smax(rho2, K.rho_min);
    rho2 = smin(rho2, K.rho_max);
    p2 = smax(p2, K.p_min);
    p2 = smin(p2, K.p_max);
    const double out[5] = {rho2, u2, v2, w2, p2};
#pragma unroll
    for (int j = 0; j < 5; ++j) q[o2 + j] = out[j];
    if (q32) {
#pragma unroll
        for (int j = 0; j < 5; ++j) q32[o2 + j] = (float)out[j];
    }
}

// Cell-face slots of one cell: device face ids, padded with the first face.
template <int MAXF>
__device__ __forceinline__ int cell_faces(const int* __restrict__ cfo, const int* __restrict__ cf, int c,
                                          int (&fs)[MAXF]) {
    const int k0 = cfo[c];
    const int nf = cfo[c + 1] - k0;
#pragma unroll
    for (int s = 0; s < MAXF; ++s) fs[s] = cf[k0 + (s < nf ? s : 0)];
    return nf;
}

// ---------------------------------------------------------------------------
// Residual: InviscidFlux + RoeFlux_noprec + LoadFlux (cal_flux.cpp:84-192,
// 682-833, 1219-1254).  res[c] starts at 0 (after ZeroResiduals) or at its
// current value (ACCUM) and receives -F for faces where c is c1 and +F where
// c is c2, in ascending face order.  Block partial sums of res^2 feed the
// residual norm of main.cpp:294-317.
template <bool ACCUM, int MAXF>
__global__ __launch_bounds__(kBlock) void k_flux(int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                 const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                                 const double* __restrict__ q, double* __restrict__ res,
                                                 double* __restrict__ part, FlowConst K) {
    const int c = xcd_block() * kBlock + (int)threadIdx.x;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < nC) {
        if (ACCUM) load5(res + (size_t)c * 5, r);
        int fs[MAXF];
        const int nf = cell_faces<MAXF>(cfo, cf, c, fs);
        int2 cc[MAXF];
        double4 g[MAXF];
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            cc[s] = fcell[fs[s]];
            g[s] = fgeo[fs[s]];
        }
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            double ql[5], qr[5], F[5];
            load5(q + (size_t)cc[s].x * 5, ql);
            load5(q + (size_t)cc[s].y * 5, qr);
            roe_flux<double>(ql, qr, g[s].x, g[s].y, g[s].z, g[s].w, K.gamm1, K.p_bar, K.alf_l, kTINYd, F);
            if (s < nf) {
                if (cc[s].x == c) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) r[j] -= F[j];
                } else {
#pragma unroll
                    for (int j = 0; j < 5; ++j) r[j] += F[j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) res[(size_t)c * 5 + j] = r[j];
    }
    double s2[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s2[j] = r[j] * r[j];
    block_sum<5>(s2, part + blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// Local time step: TimeStepNormal_new + CellIsMG + LimitTimeStep
// (compute_timestep.cpp:758-823, 1315-1338, 836-918), fused with the
// LU-SGS diagonal Diag = vol/dt + sum 0.5*A*(|u.n|+a)*lhs_omga
// (forward_lusgs.cpp:94-99, 554-595) when WITH_DIAG.  Both use the cell's
// own |u.n| + a on every face.
template <bool WITH_DIAG, int MAXF>
__global__ __launch_bounds__(kBlock) void k_timestep(int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                     const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                                     const double2* __restrict__ fdn, const double* __restrict__ q,
                                                     const double* __restrict__ vol, double* __restrict__ dt,
                                                     double* __restrict__ diag, double* __restrict__ part, FlowConst K,
                                                     double cfl, double cfl_min) {
    const int c = xcd_block() * kBlock + (int)threadIdx.x;
    double vmin = kBIGd, vmax = 0.0;
    if (c < nC) {
        double qc[5];
        load5(q + (size_t)c * 5, qc);
        const double rho = qc[0], u = qc[1], v = qc[2], w = qc[3], p = qc[4];
        int fs[MAXF];
        const int nf = cell_faces<MAXF>(cfo, cf, c, fs);
        int2 cc[MAXF];
        double4 g[MAXF];
        double2 d[MAXF];
        double pa[MAXF], pb[MAXF];
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            cc[s] = fcell[fs[s]];
            g[s] = fgeo[fs[s]];
            d[s] = fdn[fs[s]];
        }
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            pa[s] = q[(size_t)cc[s].x * 5 + 4];
            pb[s] = q[(size_t)cc[s].y * 5 + 4];
        }
        const double c2tmp = K.gam * (p + K.p_bar) / rho;
        const double a = sqrt(c2tmp);
        double dtc = kBIGd;
        int det = 1;
        double eig[MAXF];
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            const double dn = (cc[s].x == c) ? d[s].x : d[s].y;
            double vn = u * g[s].x + v * g[s].y + w * g[s].z;
            vn = fabs(vn);
            eig[s] = vn + a;
            const double dp = fabs(pb[s] - pa[s]) / (pb[s] + pa[s] + K.p_bar + K.p_bar);
            if (s < nf) {
                dtc = smin(dtc, dn / eig[s]);
                if (dp > 0.1) det = 0;
            }
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
#pragma unroll
            for (int s = 0; s < MAXF; ++s)
                if (s < nf) dg += 0.5 * g[s].w * eig[s] * K.lhs_omga;
            diag[c] = dg;
        }
    }
    block_minmax(vmin, vmax, part + (size_t)5 * gridDim.x + blockIdx.x, part + (size_t)6 * gridDim.x + blockIdx.x);
}

// Residual, local time step and LU-SGS diagonal in one pass over a cell's
// faces, for the fused iteration: the arithmetic of k_flux<false> and
// k_timestep<true>, which both read the state at the start of the
// iteration.  The cell's own state is loaded once and only the neighbour's
// is gathered per face.  With MIXED the FP32 inputs of the implicit
// operator (state snapshot, Diag) are written as well.
// Per cell: the faces' fluxes added in slot order (the cell's ascending
// face order), the time step from the smallest dn/eig, the CFL choice and
// the diagonal.  NS slots, the first nf used; oth = the face's other cell,
// g = the face's stored normal and area (c1 -> c2), first = this cell is c1.
template <int NS, bool MIXED>
__device__ __forceinline__ void flux_ts_cell(int c, const double (&qc)[5], int nf, const int (&oth)[NS],
                                             const double4 (&g)[NS], const double (&dnv)[NS], const bool (&first)[NS],
                                             const double* __restrict__ q, const double* __restrict__ vol,
                                             double* __restrict__ res, double* __restrict__ dt,
                                             double* __restrict__ diag, const FlowConst& K, double cfl, double cfl_min,
                                             float* __restrict__ q32, float* __restrict__ diag32, double (&r)[5],
                                             double& dtc) {
    const double c2tmp = K.gam * (qc[4] + K.p_bar) / qc[0];
    const double a = sqrt(c2tmp);
    dtc = kBIGd;
    int det = 1;
    double eig[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double qo[5], ql[5], qr[5], F[5];
        load5(q + (size_t)oth[s] * 5, qo);
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            ql[j] = first[s] ? qc[j] : qo[j];
            qr[j] = first[s] ? qo[j] : qc[j];
        }
        roe_flux<double>(ql, qr, g[s].x, g[s].y, g[s].z, g[s].w, K.gamm1, K.p_bar, K.alf_l, kTINYd, F);
        double vn = qc[1] * g[s].x + qc[2] * g[s].y + qc[3] * g[s].z;
        vn = fabs(vn);
        eig[s] = vn + a;
        const double dp = fabs(qr[4] - ql[4]) / (qr[4] + ql[4] + K.p_bar + K.p_bar);
        if (s < nf) {
            if (first[s]) {
#pragma unroll
                for (int j = 0; j < 5; ++j) r[j] -= F[j];
            } else {
#pragma unroll
                for (int j = 0; j < 5; ++j) r[j] += F[j];
            }
            dtc = smin(dtc, dnv[s] / eig[s]);
            if (dp > 0.1) det = 0;
        }
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) res[(size_t)c * 5 + j] = r[j];
    const double p = qc[4];
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
    double dg = vol[c] / dtc;
#pragma unroll
    for (int s = 0; s < NS; ++s)
        if (s < nf) dg += 0.5 * g[s].w * eig[s] * K.lhs_omga;
    diag[c] = dg;
    if (MIXED) {
        diag32[c] = (float)dg;
#pragma unroll
        for (int j = 0; j < 5; ++j) q32[(size_t)c * 5 + j] = (float)qc[j];
    }
}

// Residual, local time step and LU-SGS diagonal in one pass over a cell's
// faces, for the fused iteration: the arithmetic of k_flux<false> and
// k_timestep<true>, which both read the state at the start of the
// iteration.  The cell's own state is loaded once and only the neighbour's
// is gathered per face.  With MIXED the FP32 inputs of the implicit
// operator (state snapshot, Diag) are written as well.
template <int MAXF, bool MIXED>
__device__ __forceinline__ void flux_ts_body(int c, int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                             const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                             const double2* __restrict__ fdn, const double* __restrict__ q,
                                             const double* __restrict__ vol, double* __restrict__ res,
                                             double* __restrict__ dt, double* __restrict__ diag,
                                             double* __restrict__ pblk, int pstride, const FlowConst& K, double cfl,
                                             double cfl_min, float* __restrict__ q32, float* __restrict__ diag32) {
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double vmin = kBIGd, vmax = 0.0;
    if (c < nC) {
        double qc[5];
        load5(q + (size_t)c * 5, qc);
        int fs[MAXF];
        const int nf = cell_faces<MAXF>(cfo, cf, c, fs);
        int oth[MAXF];
        double4 g[MAXF];
        double dnv[MAXF];
        bool first[MAXF];
#pragma unroll
        for (int s = 0; s < MAXF; ++s) {
            const int2 cc = fcell[fs[s]];
            const double2 d = fdn[fs[s]];
            g[s] = fgeo[fs[s]];
            first[s] = (cc.x == c);  // c is the face's c1
            oth[s] = first[s] ? cc.y : cc.x;
            dnv[s] = first[s] ? d.x : d.y;
        }
        double dtc;
        flux_ts_cell<MAXF, MIXED>(c, qc, nf, oth, g, dnv, first, q, vol, res, dt, diag, K, cfl, cfl_min, q32, diag32,
                                  r, dtc);
        vmin = dtc;
        vmax = dtc;
    }
    double s2[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s2[j] = r[j] * r[j];
    block_sum<5>(s2, pblk, pstride);
    block_minmax(vmin, vmax, pblk + (size_t)5 * pstride, pblk + (size_t)6 * pstride);
}

// The same from the per-slot tables laid out by cell (neighbour, normal
// pointing away from the cell with its area, this side's distance, and the
// slots whose face has this cell second): contiguous reads, then the
// neighbours' states.  The stored normal is the slot normal negated back
// where the cell is the face's second cell, exactly.
template <int LPC, bool MIXED>
__device__ __forceinline__ void flux_ts_slot_body(int c, int nC, const int* __restrict__ snb,
                                                  const double4* __restrict__ sgeo, const double* __restrict__ sdn,
                                                  const int* __restrict__ sflip, const double* __restrict__ q,
                                                  const double* __restrict__ vol, double* __restrict__ res,
                                                  double* __restrict__ dt, double* __restrict__ diag,
                                                  double* __restrict__ pblk, int pstride, const FlowConst& K, double cfl,
                                                  double cfl_min, float* __restrict__ q32,
                                                  float* __restrict__ diag32) {
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double vmin = kBIGd, vmax = 0.0;
    if (c < nC) {
        double qc[5];
        load5(q + (size_t)c * 5, qc);
        const unsigned sf = (unsigned)sflip[c];
        const int nf = (int)((sf >> 8) & 15u);
        int oth[LPC];
        double4 g[LPC];
        double dnv[LPC];
        bool first[LPC];
#pragma unroll
        for (int s = 0; s < LPC; ++s) {
            const size_t i = (size_t)c * LPC + s;
            const double4 v = sgeo[i];
            first[s] = ((sf >> s) & 1u) == 0;
            g[s] = first[s] ? v : make_double4(-v.x, -v.y, -v.z, v.w);
            oth[s] = snb[i];
            dnv[s] = sdn[i];
        }
        double dtc;
        flux_ts_cell<LPC, MIXED>(c, qc, nf, oth, g, dnv, first, q, vol, res, dt, diag, K, cfl, cfl_min, q32, diag32,
                                 r, dtc);
        vmin = dtc;
        vmax = dtc;
    }
    double s2[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s2[j] = r[j] * r[j];
    block_sum<5>(s2, pblk, pstride);
    block_minmax(vmin, vmax, pblk + (size_t)5 * pstride, pblk + (size_t)6 * pstride);
}

template <int LPC, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_flux_ts_slot(int nC, const int* __restrict__ snb,
                                                         const double4* __restrict__ sgeo,
                                                         const double* __restrict__ sdn, const int* __restrict__ sflip,
                                                         const double* __restrict__ q, const double* __restrict__ vol,
                                                         double* __restrict__ res, double* __restrict__ dt,
                                                         double* __restrict__ diag, double* __restrict__ part,
                                                         FlowConst K, double cfl, double cfl_min,
                                                         float* __restrict__ q32, float* __restrict__ diag32) {
    flux_ts_slot_body<LPC, MIXED>(xcd_block() * kBlock + (int)threadIdx.x, nC, snb, sgeo, sdn, sflip, q, vol, res,
                                  dt, diag, part + blockIdx.x, (int)gridDim.x, K, cfl, cfl_min, q32, diag32);
}

template <int MAXF, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_flux_ts(int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                    const int2* __restrict__ fcell, const double4* __restrict__ fgeo,
                                                    const double2* __restrict__ fdn, const double* __restrict__ q,
                                                    const double* __restrict__ vol, double* __restrict__ res,
                                                    double* __restrict__ dt, double* __restrict__ diag,
                                                    double* __restrict__ part, FlowConst K, double cfl,
                                                    double cfl_min, float* __restrict__ q32,
                                                    float* __restrict__ diag32) {
    flux_ts_body<MAXF, MIXED>(xcd_block() * kBlock + (int)threadIdx.x, nC, cfo, cf, fcell, fgeo, fdn, q, vol, res, dt,
                              diag, part + blockIdx.x, (int)gridDim.x, K, cfl, cfl_min, q32, diag32);
}

// Diag alone (forward_lusgs.cpp:94-99 + CalDiagLUSGS :554-595), from dt.
// With `run`, only cells whose dt exceeds ratio_dtmax*dt_min are clamped
// and recomputed (compute_timestep.cpp:929-945), and only when the
// reduction found the ratio exceeded.
template <int MAXF>
__device__ __forceinline__ void diag_body(int c, int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                          const double4* __restrict__ fgeo, const double* __restrict__ q,
                                          const double* __restrict__ vol, double* __restrict__ dt,
                                          double* __restrict__ diag, const double* __restrict__ run,
                                          double ratio_dtmax, const FlowConst& K, float* __restrict__ diag32) {
    if (c >= nC) return;
    double dtc = dt[c];
    if (run) {
        if (run[2] == 0.0) return;
        const double lim = ratio_dtmax * run[1];
        if (!(dtc > lim)) return;
        dtc = lim;
        dt[c] = dtc;
    }
    double qc[5];
    load5(q + (size_t)c * 5, qc);
    int fs[MAXF];
    const int nf = cell_faces<MAXF>(cfo, cf, c, fs);
    const double a = sqrt(K.gam * (qc[4] + K.p_bar) / qc[0]);
    double dg = vol[c] / dtc;
#pragma unroll
    for (int s = 0; s < MAXF; ++s) {
        const double4 g = fgeo[fs[s]];
        double vn = qc[1] * g.x + qc[2] * g.y + qc[3] * g.z;
        vn = fabs(vn);
        if (s < nf) dg += 0.5 * g.w * (vn + a) * K.lhs_omga;
    }
    diag[c] = dg;
    if (diag32) diag32[c] = (float)dg;
}
template <int MAXF>
__global__ __launch_bounds__(kBlock) void k_diag(int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                 const double4* __restrict__ fgeo, const double* __restrict__ q,
                                                 const double* __restrict__ vol, double* __restrict__ dt,
                                                 double* __restrict__ diag, const double* __restrict__ run,
                                                 double ratio_dtmax, FlowConst K, float* __restrict__ diag32) {
    diag_body<MAXF>(xcd_block() * kBlock + (int)threadIdx.x, nC, cfo, cf, fgeo, q, vol, dt, diag, run, ratio_dtmax, K,
                    diag32);
}

// The ratio_dtmax clamp alone (compute_timestep.cpp:929-945) over a small
// grid: every workgroup reads the reduction's flag first and returns when
// the clamp is off, which is the rule at the reference's ratio_dtmax.
template <int MAXF>
__global__ __launch_bounds__(kBlock) void k_clamp(int nC, const int* __restrict__ cfo, const int* __restrict__ cf,
                                                  const double4* __restrict__ fgeo, const double* __restrict__ q,
                                                  const double* __restrict__ vol, double* __restrict__ dt,
                                                  double* __restrict__ diag, const double* __restrict__ run,
                                                  double ratio_dtmax, FlowConst K, float* __restrict__ diag32) {
    if (run[2] == 0.0) return;
    for (int c = (int)(blockIdx.x * kBlock + threadIdx.x); c < nC; c += (int)(gridDim.x * kBlock))
        diag_body<MAXF>(c, nC, cfo, cf, fgeo, q, vol, dt, diag, run, ratio_dtmax, K, diag32);
}

// FP32 operator inputs of a mixed-precision context from the FP64 state and
// Diag (interior cells; the ghost part of the snapshot never changes).
__global__ __launch_bounds__(kBlock) void k_mixed_inputs(int nC, const double* __restrict__ q,
                                                         const double* __restrict__ diag, float* __restrict__ q32,
                                                         float* __restrict__ diag32) {
    const int c = blockIdx.x * kBlock + (int)threadIdx.x;
    if (c >= nC) return;
#pragma unroll
    for (int j = 0; j < 5; ++j) q32[(size_t)c * 5 + j] = (float)q[(size_t)c * 5 + j];
    diag32[c] = (float)diag[c];
}

// ---------------------------------------------------------------------------
// Reductions.  mode bits: 1 = residual norms -> run[3..7], 2 = dt extrema
// (running max/min of compute_timestep.cpp:914-918, clamp flag), 4 = write
// history record `slot` from run[].
// One workgroup of kReduceBlock threads reads each block's partials once
// (norms and extrema together) and sums them in a fixed tree.  The partials
// are stored value-major, part[j * nblk + block] for the five squared
// norms (j = 0..4), dt min (5) and dt max (6), so that every load of a wave
// reads 64 consecutive doubles: with one 64-byte record per block, a wave's
// load touched 32 cache lines, and a 6.0M-cell reduction took 57 us in the
// single workgroup's load unit.
constexpr int kReduceBlock = 1024;

// The second half of a reduction: the block tree over the threads' folds
// (s: squared-norm sums, a: dt min, b: dt max), the running extrema, the
// clamp flag and the history record.
__device__ __forceinline__ void reduce_finish(double (&s)[5], double a, double b, double* __restrict__ run,
                                              double* __restrict__ hist, int slot, int mode, double ratio_dtmax) {
    constexpr int NW = kReduceBlock / 64;
    if (mode & 1) {
        __shared__ double tot[5];
        block_sum<5, NW>(s, tot);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int j = 0; j < 5; ++j) run[3 + j] = sqrt(tot[j]);
    }
    if (mode & 2) {
        __shared__ double mn, mx;
        block_minmax<NW>(a, b, &mn, &mx);
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
    if (threadIdx.x == 0 && (mode & 4)) {
        double* h = hist + (size_t)slot * 7;
        for (int j = 0; j < 5; ++j) h[j] = run[3 + j];
        h[5] = run[0];
        h[6] = run[1];
    }
}

__device__ __forceinline__ void reduce_body(int nblk, const double* __restrict__ part, double* __restrict__ run,
                                            double* __restrict__ hist, int slot, int mode, double ratio_dtmax) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double a = kBIGd, b = 0.0;
    // each thread folds records k, k+1024, ... in that order; the records
    // of a group of kReduceBatch are loaded together before they are added
    // (a group past the end loads a valid record in place of the missing
    // ones and skips them), so the thread waits for one round trip per
    // group, not per record
    constexpr int kReduceBatch = 7;  // 8 spills registers
    for (int k = threadIdx.x; k < nblk; k += kReduceBatch * kReduceBlock) {
        double r[kReduceBatch][7];
#pragma unroll
        for (int i = 0; i < kReduceBatch; ++i) {
            const int e = min(k + i * kReduceBlock, nblk - 1);
#pragma unroll
            for (int j = 0; j < 7; ++j) r[i][j] = part[(size_t)j * nblk + e];
        }
#pragma unroll
        for (int i = 0; i < kReduceBatch; ++i) {
            if (k + i * kReduceBlock >= nblk) break;
            if (mode & 1) {
#pragma unroll
                for (int j = 0; j < 5; ++j) s[j] += r[i][j];
            }
            if (mode & 2) {
                a = smin(a, r[i][5]);
                b = smax(b, r[i][6]);
            }
        }
    }
    reduce_finish(s, a, b, run, hist, slot, mode, ratio_dtmax);
}
__global__ __launch_bounds__(kReduceBlock) void k_reduce(int nblk, const double* __restrict__ part, double* __restrict__ run,
                                                   double* __restrict__ hist, int slot, int mode, double ratio_dtmax) {
    reduce_body(nblk, part, run, hist, slot, mode, ratio_dtmax);
}

// The same reduction in two launches, for many blocks.  One workgroup
// reading all the records is bound by the misses one CU keeps in flight
// (42 us for the 23,438 records of 6.0M cells, whatever the batching).
// k_reduce_fold does the threads' folds on many CUs: the fold of thread t
// of k_reduce (records t, t+1024, ... in that order) is shared by
// kFoldLanes consecutive lanes, each loading up to kFoldRecs of its
// records at once; the running values pass from lane to lane in record
// order, so every addition is k_reduce's, in k_reduce's order.  The folds
// go to fold[j * 1024 + t] (j: 5 sums, dt min, dt max) and k_reduce_tree
// finishes them exactly as k_reduce does, so both give the same bits.
constexpr int kFoldLanes = 8, kFoldRecs = 4;
constexpr int kFoldThreads = 256;
__global__ __launch_bounds__(kFoldThreads) void k_reduce_fold(int nblk, const double* __restrict__ part,
                                                              double* __restrict__ fold) {
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int t = g / kFoldLanes;  // thread of k_reduce whose fold this is
    const int sub = g % kFoldLanes;
    const int lane = (int)threadIdx.x & 63;
    const int src = lane - sub + (sub + kFoldLanes - 1) % kFoldLanes;  // the previous lane of this fold
    const int cnt = t < nblk ? (nblk - 1 - t) / kReduceBlock + 1 : 0;  // records of thread t
    const int maxcnt = (nblk + kReduceBlock - 1) / kReduceBlock;       // the same bound for every lane
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double a = kBIGd, b = 0.0;
    for (int base = 0; base < maxcnt; base += kFoldLanes * kFoldRecs) {
        double r[kFoldRecs][7];
#pragma unroll
        for (int m = 0; m < kFoldRecs; ++m) {
            const int i = base + sub * kFoldRecs + m;
            const int e = min(t + i * kReduceBlock, nblk - 1);
#pragma unroll
            for (int j = 0; j < 7; ++j) r[m][j] = part[(size_t)j * nblk + e];
        }
        // lane `step` adds its records to the running values, which then
        // move on to the next lane (from the last lane back to the first,
        // which starts the next round)
        for (int step = 0; step < kFoldLanes; ++step) {
            if (sub == step) {
#pragma unroll
                for (int m = 0; m < kFoldRecs; ++m) {
                    if (base + sub * kFoldRecs + m >= cnt) break;
#pragma unroll
                    for (int j = 0; j < 5; ++j) s[j] += r[m][j];
                    a = smin(a, r[m][5]);
                    b = smax(b, r[m][6]);
                }
            }
            double ts[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) ts[j] = __shfl(s[j], src, 64);
            const double ta = __shfl(a, src, 64), tb = __shfl(b, src, 64);
            if (sub == (step + 1) % kFoldLanes) {
#pragma unroll
                for (int j = 0; j < 5; ++j) s[j] = ts[j];
                a = ta;
                b = tb;
            }
        }
    }
    if (sub == 0 && t < kReduceBlock) {  // the complete fold is back at the first lane
#pragma unroll
        for (int j = 0; j < 5; ++j) fold[j * kReduceBlock + t] = s[j];
        fold[5 * kReduceBlock + t] = a;
        fold[6 * kReduceBlock + t] = b;
    }
}
__global__ __launch_bounds__(kReduceBlock) void k_reduce_tree(const double* __restrict__ fold, double* __restrict__ run,
                                                        double* __restrict__ hist, int slot, int mode,
                                                        double ratio_dtmax) {
    const int t = (int)threadIdx.x;
    double s[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) s[j] = (mode & 1) ? fold[j * kReduceBlock + t] : 0.0;
    const double a = (mode & 2) ? fold[5 * kReduceBlock + t] : kBIGd;
    const double b = (mode & 2) ? fold[6 * kReduceBlock + t] : 0.0;
    reduce_finish(s, a, b, run, hist, slot, mode, ratio_dtmax);
}

// ---------------------------------------------------------------------------
// Node values (CompNodeVar3D_dist, cal_gradient.cpp:1922-2372): each node
// sums its weighted contributions in the reference's order, then divides by
// WeightNode + TINY.
// Entries are stored per wave of 64 consecutive nodes, entry-major
// (entry k of the wave's lane i at ell_off[w] + 64k + i, padded to the
// wave's largest count), so every entry load of a wave is one coalesced
// 512-byte row; each node still adds its own entries in its order.
template <typename R>
__device__ __forceinline__ void node_body(int p, int w, int nN, const int* __restrict__ nlo,
                                          const int* __restrict__ ell_off, const int2* __restrict__ ell_ab,
                                          const R* __restrict__ ell_w, const double* __restrict__ nden,
                                          const double* __restrict__ q, R* __restrict__ qn) {
    const int lane = (int)threadIdx.x & 63;
    const int e0 = ell_off[w];
    const int K = (ell_off[w + 1] - e0) / 64;  // wave-uniform trip count
    const int cnt = p < nN ? nlo[p + 1] - nlo[p] : 0;
    R acc[5] = {R(0), R(0), R(0), R(0), R(0)};
    // entries in groups of U: the gathers of a group are issued together,
    // then added in entry order
    constexpr int U = 6;
    for (int k = 0; k < K; k += U) {
        int2 ab[U];
        R wt[U], qa[U][5];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const int kk = (k + i < K) ? k + i : k;
            ab[i] = ell_ab[(size_t)e0 + (size_t)kk * 64 + lane];
            wt[i] = ell_w[(size_t)e0 + (size_t)kk * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < U; ++i)
#pragma unroll
            for (int j = 0; j < 5; ++j) qa[i][j] = (R)q[(size_t)ab[i].x * 5 + j];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            if (k + i >= cnt) break;
            if (ab[i].y < 0) {
#pragma unroll
                for (int j = 0; j < 5; ++j) acc[j] += qa[i][j] * wt[i];
            } else {
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    R fq = qa[i][j] + (R)q[(size_t)ab[i].y * 5 + j];
                    fq *= R(0.5);
                    acc[j] += fq * wt[i];
                }
            }
        }
    }
    if (p >= nN) return;
    const R d = (R)nden[p];
#pragma unroll
    for (int j = 0; j < 5; ++j) qn[(size_t)p * 5 + j] = acc[j] / d;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_node(int nN, const int* __restrict__ nlo, const int* __restrict__ ell_off,
                                                 const int2* __restrict__ ell_ab, const R* __restrict__ ell_w,
                                                 const double* __restrict__ nden, const double* __restrict__ q,
                                                 R* __restrict__ qn) {
    const int lb = xcd_block();
    node_body<R>(lb * kBlock + (int)threadIdx.x, lb * (kBlock / 64) + (int)threadIdx.x / 64, nN, nlo, ell_off, ell_ab,
                 ell_w, nden, q, qn);
}

// Symmetry-plane velocity projection of the node values (cal_gradient.cpp:2332-2360).
template <typename R>
__global__ void k_node_symm(int nsym, const int* __restrict__ sym_node, const double* __restrict__ sym_n,
                            R* __restrict__ qn) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsym) return;
    const size_t p = (size_t)sym_node[s] * 5;
    const R xs = (R)sym_n[3 * s], ys = (R)sym_n[3 * s + 1], zs = (R)sym_n[3 * s + 2];
    const R un = qn[p + 1], vn = qn[p + 2], wn = qn[p + 3];
    const R vv = un * xs + vn * ys + wn * zs;
    qn[p + 1] = un - vv * xs;
    qn[p + 2] = vn - vv * ys;
    qn[p + 3] = wn - vv * zs;
}

// Green-Gauss gradient (CompGradientQ_Gauss_Node, cal_gradient.cpp:1740-1795,
// 1889-1897): face values from the node values (interior faces, SYMM and
// INTERFACE boundary faces) or the face-centre average of cell and ghost;
// outputs in the reference's per-variable blocks of nC+nB.
template <int MAXF, typename R>
__device__ __forceinline__ void grad_body(int c, int nC, int n, const int* __restrict__ cfo,
                                          const int* __restrict__ cf, const int2* __restrict__ fcell,
                                          const typename Real<R>::vec4* __restrict__ fgeo,
                                          const int* __restrict__ fno, const int* __restrict__ fn,
                                          const int* __restrict__ btype, const double* __restrict__ q,
                                          const R* __restrict__ qn, const double* __restrict__ vol,
                                          R* __restrict__ gx, R* __restrict__ gy, R* __restrict__ gz) {
    if (c >= nC) return;
    R ax[5] = {R(0), R(0), R(0), R(0), R(0)}, ay[5] = {R(0), R(0), R(0), R(0), R(0)},
      az[5] = {R(0), R(0), R(0), R(0), R(0)};
    int fs[MAXF];
    const int nf = cell_faces<MAXF>(cfo, cf, c, fs);
    int2 cc[MAXF];
    typename Real<R>::vec4 g[MAXF];
    // tetrahedral meshes: every face is a triangle and its nodes sit at 3f
    int tn[MAXF == 4 ? 12 : 1];
#pragma unroll
    for (int s = 0; s < MAXF; ++s) {
        cc[s] = fcell[fs[s]];
        g[s] = fgeo[fs[s]];
        if constexpr (MAXF == 4) {
#pragma unroll
            for (int i = 0; i < 3; ++i) tn[3 * s + i] = fn[3 * fs[s] + i];
        }
    }
#pragma unroll
    for (int s = 0; s < MAXF; ++s) {
        if (s >= nf) break;
        const int f = fs[s];
        R qs[5];
        bool node_mean = true;
        if (cc[s].y >= nC) {
            const int t = btype[cc[s].y - nC];
            node_mean = (t == BC_INTERFACE || t == BC_SYMM);
        }
        if (node_mean) {
            R sum[5] = {R(0), R(0), R(0), R(0), R(0)};
            R cnt;
            if constexpr (MAXF == 4) {
                R v0[5], v1[5], v2[5];
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    v0[j] = qn[(size_t)tn[3 * s] * 5 + j];
                    v1[j] = qn[(size_t)tn[3 * s + 1] * 5 + j];
                    v2[j] = qn[(size_t)tn[3 * s + 2] * 5 + j];
                }
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    sum[j] += v0[j];
                    sum[j] += v1[j];
                    sum[j] += v2[j];
                }
                cnt = R(3.0);
            } else {
                const int j0 = fno[f], j1 = fno[f + 1];
                cnt = (R)(j1 - j0);
                for (int m = j0; m < j1; ++m) {
#pragma unroll
                    for (int j = 0; j < 5; ++j) sum[j] += qn[(size_t)fn[m] * 5 + j];
                }
            }
#pragma unroll
            for (int j = 0; j < 5; ++j) qs[j] = sum[j] / cnt;
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j)
                qs[j] = R(0.5) * ((R)q[(size_t)cc[s].x * 5 + j] + (R)q[(size_t)cc[s].y * 5 + j]);
        }
        const bool minus = (cc[s].y < nC) && (cc[s].y == c);
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const R sa = qs[j] * g[s].w;
            const R tx = sa * g[s].x, ty = sa * g[s].y, tz = sa * g[s].z;
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
    const R vc = (R)vol[c];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        gx[(size_t)j * n + c] = ax[j] / vc;
        gy[(size_t)j * n + c] = ay[j] / vc;
        gz[(size_t)j * n + c] = az[j] / vc;
    }
}

// The same gradient for tetrahedral cells, from per-slot tables laid out
// by cell: the neighbour across each face (snb), the face's unit normal
// pointing away from the cell with its area (sgeo), the face's three nodes
// in the face's own order (snode), and in bits 24-27 of depmask the
// boundary slots that take the cell/ghost average.  A cell reads 4 + 16 +
// 12 contiguous values and then gathers its node values, instead of
// walking face list -> face -> nodes.  Subtracting q_f*A*n for a face whose
// second cell this is equals adding q_f*A*(-n) exactly, so the sums are
// the same bits as grad_body's.
template <typename R>
__device__ __forceinline__ void grad_tet_body(int c, int nC, int n, const int* __restrict__ snb,
                                              const typename Real<R>::vec4* __restrict__ sgeo,
                                              const int* __restrict__ snode, const int* __restrict__ depmask,
                                              const double* __restrict__ q, const R* __restrict__ qn,
                                              const double* __restrict__ vol, R* __restrict__ gx,
                                              R* __restrict__ gy, R* __restrict__ gz) {
    if (c >= nC) return;
    const unsigned fca = (unsigned)depmask[c] >> 24;
    const int4 nb4 = reinterpret_cast<const int4*>(snb)[c];
    const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    typename Real<R>::vec4 g[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) g[s] = sgeo[(size_t)c * 4 + s];
    const int4* sn = reinterpret_cast<const int4*>(snode + (size_t)c * 12);
    const int4 t0 = sn[0], t1 = sn[1], t2 = sn[2];
    const int tn[12] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w, t2.x, t2.y, t2.z, t2.w};
    R ax[5] = {R(0), R(0), R(0), R(0), R(0)}, ay[5] = {R(0), R(0), R(0), R(0), R(0)},
      az[5] = {R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        R qs[5];
        if ((fca >> s) & 1u) {
#pragma unroll
            for (int j = 0; j < 5; ++j)
                qs[j] = R(0.5) * ((R)q[(size_t)c * 5 + j] + (R)q[(size_t)nbs[s] * 5 + j]);
        } else {
            R sum[5] = {R(0), R(0), R(0), R(0), R(0)};
            R v0[5], v1[5], v2[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                v0[j] = qn[(size_t)tn[3 * s] * 5 + j];
                v1[j] = qn[(size_t)tn[3 * s + 1] * 5 + j];
                v2[j] = qn[(size_t)tn[3 * s + 2] * 5 + j];
            }
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                sum[j] += v0[j];
                sum[j] += v1[j];
                sum[j] += v2[j];
            }
#pragma unroll
            for (int j = 0; j < 5; ++j) qs[j] = sum[j] / R(3.0);
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const R sa = qs[j] * g[s].w;
            ax[j] += sa * g[s].x;
            ay[j] += sa * g[s].y;
            az[j] += sa * g[s].z;
        }
    }
    const R vc = (R)vol[c];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        // nothing in the iteration reads the gradient: streaming stores keep
        // its 120 bytes per cell from displacing the sweeps' data in the
        // caches (it runs beside the sweeps' thin tails)
        __builtin_nontemporal_store(ax[j] / vc, &gx[(size_t)j * n + c]);
        __builtin_nontemporal_store(ay[j] / vc, &gy[(size_t)j * n + c]);
        __builtin_nontemporal_store(az[j] / vc, &gz[(size_t)j * n + c]);
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_grad_tet(int nC, int n, const int* __restrict__ snb,
                                                     const typename Real<R>::vec4* __restrict__ sgeo,
                                                     const int* __restrict__ snode, const int* __restrict__ depmask,
                                                     const double* __restrict__ q, const R* __restrict__ qn,
                                                     const double* __restrict__ vol, R* __restrict__ gx,
                                                     R* __restrict__ gy, R* __restrict__ gz) {
    grad_tet_body<R>(xcd_block() * kBlock + (int)threadIdx.x, nC, n, snb, sgeo, snode, depmask, q, qn, vol, gx, gy,
                     gz);
}

// k_grad_tet on a small grid whose workgroups claim blocks of kBlock cells:
// workgroup b takes the next unclaimed block of XCD x = b mod 8's span
// (blocks x*span .. (x+1)*span-1, span = nblk / 8, the blocks xcd_block()
// puts on that XCD) from counter x, until the span is done.  The next
// block is claimed while the current one is worked on.  The last workgroup
// to find its span done sets the counter back to zero for the next launch
// (each of the G/8 workgroups of an XCD ends with exactly one claim past
// the span, so the claim returning span + G/8 - 1 is the last).  nblk and
// the grid G are multiples of 8; the counters are 64 ints apart and start
// at zero.
//
// Launched on one workgroup per CU on the context's second stream, beside
// the sweeps' thin tails (6.0M cells, reference order): the gradient takes
// as long as on the full grid (23,440 workgroups), but its waves crowd the
// tails' CUs less (FP64 iteration 3.63 -> 3.57 ms; mixed 3.04 -> 2.99).
// Beside wide sweep levels (the shallower 2/128 order, no thin tail): FP64
// 3.01 -> 2.95 ms, mixed 2.31 -> 2.28; a fixed share per workgroup instead
// of claims left the late-starting workgroups' shares to the end there (the
// gradient 1.40 -> 2.3 ms).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_grad_tet_claim(int nblk, int nC, int n, const int* __restrict__ snb,
                                                           const typename Real<R>::vec4* __restrict__ sgeo,
                                                           const int* __restrict__ snode,
                                                           const int* __restrict__ depmask,
                                                           const double* __restrict__ q, const R* __restrict__ qn,
                                                           const double* __restrict__ vol, R* __restrict__ gx,
                                                           R* __restrict__ gy, R* __restrict__ gz, int* counters) {
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
        grad_tet_body<R>((x * span + k) * kBlock + (int)threadIdx.x, nC, n, snb, sgeo, snode, depmask, q, qn, vol,
                         gx, gy, gz);
        if (threadIdx.x == 0) claimed = next;
        __syncthreads();
        k = claimed;
    }
    if (threadIdx.x == 0 && k == span + per - 1) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The gradient of one iteration and the residual / time step / Diag pass of
// the next, on tetrahedral meshes, in one launch.  One launch instead of
// two lets each wave overlap the gradient's node gathers with the flux
// pass's neighbour gathers (the HBM bytes stay those of the two passes: the
// gathers in between evict the shared slot data).  The gradient feeds
// nothing the next iteration computes, and the two write disjoint arrays,
// so every value is the same as with two launches.
template <bool MIXED, typename R>
__global__ __launch_bounds__(kBlock) void k_grad_flux_ts(int nC, int n, const int* __restrict__ snb,
                                                        const typename Real<R>::vec4* __restrict__ sgeo_r,
                                                        const int* __restrict__ snode, const int* __restrict__ depmask,
                                                        const double* __restrict__ q, const R* __restrict__ qn,
                                                        const double* __restrict__ vol, R* __restrict__ gx,
                                                        R* __restrict__ gy, R* __restrict__ gz,
                                                        const double4* __restrict__ sgeo, const double* __restrict__ sdn,
                                                        const int* __restrict__ sflip, double* __restrict__ res,
                                                        double* __restrict__ dt, double* __restrict__ diag,
                                                        double* __restrict__ part, FlowConst K, double cfl,
                                                        double cfl_min, float* __restrict__ q32,
                                                        float* __restrict__ diag32) {
    const int c = xcd_block() * kBlock + (int)threadIdx.x;
    grad_tet_body<R>(c, nC, n, snb, sgeo_r, snode, depmask, q, qn, vol, gx, gy, gz);  // no-op past nC
    flux_ts_slot_body<4, MIXED>(c, nC, snb, sgeo, sdn, sflip, q, vol, res, dt, diag,
                                part + blockIdx.x, (int)gridDim.x, K, cfl, cfl_min, q32, diag32);
}

template <int MAXF, typename R>
__global__ __launch_bounds__(kBlock) void k_grad(int nC, int n, const int* __restrict__ cfo,
                                                 const int* __restrict__ cf, const int2* __restrict__ fcell,
                                                 const typename Real<R>::vec4* __restrict__ fgeo,
                                                 const int* __restrict__ fno, const int* __restrict__ fn,
                                                 const int* __restrict__ btype, const double* __restrict__ q,
                                                 const R* __restrict__ qn, const double* __restrict__ vol,
                                                 R* __restrict__ gx, R* __restrict__ gy, R* __restrict__ gz) {
    grad_body<MAXF, R>(xcd_block() * kBlock + (int)threadIdx.x, nC, n, cfo, cf, fcell, fgeo, fno, fn, btype, q, qn,
                       vol, gx, gy, gz);
}
