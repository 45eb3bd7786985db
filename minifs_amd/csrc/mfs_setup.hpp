// ComputeMetrics on the GPU (grid_polyhedra_auxiliary_functions.cpp:68-808),
// included by mfs_device.hip inside namespace mfs::dev.
//
// The reference computes the metrics with serial face loops that scatter
// into cells.  Here every scatter becomes a gather: a cell walks its faces
// in ascending face id (the order in which the serial loops reach it: the
// boundary faces come first and both loops ascend), and a face's own work
// (node sums, the fixed-point centroid passes, normals) is one thread's
// loop over its nodes in the file's order.  Each value is therefore formed
// by the same operations in the same order as on the host, and the results
// are bitwise those of compute_metrics (mfs_grid.cpp) built with
// -ffp-contract=off.  Division and sqrt are IEEE double on both sides; the
// one pow of the reference (CorrectCellCentroid's 10^-15) is evaluated on
// the host and passed in.

constexpr double kMetricTiny = 1.e-40;  // TINY of the reference (constant.h)

// cell -> faces: counts, then a fill in arbitrary order, then each cell's
// list sorted ascending (the order the serial loops visit a cell's faces)
__global__ __launch_bounds__(kBlock) void k_c2f_count(int nF, int nB, const int* __restrict__ f2c, int* cnt) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= nF) return;
    atomicAdd(cnt + f2c[2 * i], 1);
    if (i >= nB) atomicAdd(cnt + f2c[2 * i + 1], 1);
}

__global__ __launch_bounds__(kBlock) void k_c2f_fill(int nF, int nB, const int* __restrict__ f2c,
                                                    const int* __restrict__ cfo, int* cursor, int* cf) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= nF) return;
    const int c1 = f2c[2 * i];
    cf[cfo[c1] + atomicAdd(cursor + c1, 1)] = i;
    if (i >= nB) {
        const int c2 = f2c[2 * i + 1];
        cf[cfo[c2] + atomicAdd(cursor + c2, 1)] = i;
    }
}

__global__ __launch_bounds__(kBlock) void k_c2f_sort(int nC, const int* __restrict__ cfo, int* cf) {
    const int c = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (c >= nC) return;
    const int b = cfo[c], e = cfo[c + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = cf[k];
        int j = k - 1;
        while (j >= b && cf[j] > v) {
            cf[j + 1] = cf[j];
            --j;
        }
        cf[j + 1] = v;
    }
}

struct MetricArgs {
    int nC, nB, nF;
    const int* fno;  // nF+1: first entry of each face in f2n
    const int* f2n;
    const int* f2c;
    const double *x, *y, *z;
    const int *cfo, *cf;  // cell -> faces, ascending
    double *xfc, *yfc, *zfc, *xfn, *yfn, *zfn, *area;  // nF
    double *xcc, *ycc, *zcc, *vol;                    // nC+nB
    double *xcct, *ycct, *zcct;                       // nC: volume centroid
    int* small;                                       // cells with a too small volume
};

// FaceCellCenterbyAverage (:271-300): face centre = mean of its nodes
__global__ __launch_bounds__(kBlock) void k_face_avg(MetricArgs A) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= A.nF) return;
    const int b = A.fno[i], n = A.fno[i + 1] - b;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = 0; j < n; ++j) {
        const int p = A.f2n[b + j];
        sx += A.x[p];
        sy += A.y[p];
        sz += A.z[p];
    }
    A.xfc[i] = sx / n;
    A.yfc[i] = sy / n;
    A.zfc[i] = sz / n;
}

// FaceCellCenterbyAverage (:300-355): cell centre = mean of its face centres
__global__ __launch_bounds__(kBlock) void k_cell_avg(MetricArgs A) {
    const int c = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (c >= A.nC) return;
    const int b = A.cfo[c], n = A.cfo[c + 1] - b;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < n; ++k) {
        const int i = A.cf[b + k];
        sx += A.xfc[i];
        sy += A.yfc[i];
        sz += A.zfc[i];
    }
    A.xcc[c] = sx / n;
    A.ycc[c] = sy / n;
    A.zcc[c] = sz / n;
}

// FaceAreaNormalCentroid_cycle (:68-266): four fixed-point passes on the
// face centroid over the triangles (node j, node j+1, centre), then the
// normal, area and centroid; a degenerate face takes the direction from
// its first cell's centre.
__global__ __launch_bounds__(kBlock) void k_face_geom(MetricArgs A) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= A.nF) return;
    const int b = A.fno[i], n = A.fno[i + 1] - b;
    const int* const fn = A.f2n + b;
    double xf = A.xfc[i], yf = A.yfc[i], zf = A.zfc[i];
    for (int k = 0; k < 4; ++k) {
        double ar = 0.0, xs = 0.0, ys = 0.0, zs = 0.0;
        const int p1 = fn[0];
        for (int j = 0; j < n; ++j) {
            const int p2 = fn[j];
            const int p3 = (j == n - 1) ? p1 : fn[j + 1];
            const double x21 = A.x[p2] - xf, y21 = A.y[p2] - yf, z21 = A.z[p2] - zf;
            const double x31 = A.x[p3] - xf, y31 = A.y[p3] - yf, z31 = A.z[p3] - zf;
            const double xn = y21 * z31 - y31 * z21;
            const double yn = z21 * x31 - z31 * x21;
            const double zn = x21 * y31 - x31 * y21;
            const double art = sqrt(xn * xn + yn * yn + zn * zn);
            const double xt = A.x[p2] + A.x[p3] + xf;
            const double yt = A.y[p2] + A.y[p3] + yf;
            const double zt = A.z[p2] + A.z[p3] + zf;
            ar += art;
            xs += art * xt;
            ys += art * yt;
            zs += art * zt;
        }
        if (ar > kMetricTiny) {
            const double art = 1.0 / ar / 3.0;
            xf = xs * art;
            yf = ys * art;
            zf = zs * art;
        }
    }
    double ar = 0.0, xs = 0.0, ys = 0.0, zs = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
    const int p1 = fn[0];
    for (int j = 0; j < n; ++j) {
        const int p2 = fn[j];
        const int p3 = (j == n - 1) ? p1 : fn[j + 1];
        const double x21 = A.x[p2] - xf, y21 = A.y[p2] - yf, z21 = A.z[p2] - zf;
        const double x31 = A.x[p3] - xf, y31 = A.y[p3] - yf, z31 = A.z[p3] - zf;
        const double xn = y21 * z31 - y31 * z21;
        const double yn = z21 * x31 - z31 * x21;
        const double zn = x21 * y31 - x31 * y21;
        const double art = sqrt(xn * xn + yn * yn + zn * zn);
        const double xt = A.x[p2] + A.x[p3] + xf;
        const double yt = A.y[p2] + A.y[p3] + yf;
        const double zt = A.z[p2] + A.z[p3] + zf;
        ar += art;
        xs += art * xt;
        ys += art * yt;
        zs += art * zt;
        nx += xn;
        ny += yn;
        nz += zn;
    }
    if (ar > kMetricTiny) {
        const double art = 1.0 / ar / 3.0;
        xf = xs * art;
        yf = ys * art;
        zf = zs * art;
    }
    const double art = sqrt(nx * nx + ny * ny + nz * nz);
    A.area[i] = art * 0.5;
    if (art > kMetricTiny) {
        nx /= art;
        ny /= art;
        nz /= art;
    } else {
        const int c1 = A.f2c[2 * i];
        const double xn = xf - A.xcc[c1], yn = yf - A.ycc[c1], zn = zf - A.zcc[c1];
        const double a2 = sqrt(xn * xn + yn * yn + zn * zn);
        nx = xn / a2;
        ny = yn / a2;
        nz = zn / a2;
    }
    A.xfc[i] = xf;
    A.yfc[i] = yf;
    A.zfc[i] = zf;
    A.xfn[i] = nx;
    A.yfn[i] = ny;
    A.zfn[i] = nz;
}

// CellVolCentroid (:357-536): the volume and volume centroid of each cell
// from the pyramids (face triangle, cell centre); the face's first cell
// takes the negated volume.
__global__ __launch_bounds__(kBlock) void k_cell_vol(MetricArgs A) {
    const int c = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (c >= A.nC) return;
    const double xc = A.xcc[c], yc = A.ycc[c], zc = A.zcc[c];
    double v = 0.0, xs = 0.0, ys = 0.0, zs = 0.0;
    for (int k = A.cfo[c]; k < A.cfo[c + 1]; ++k) {
        const int i = A.cf[k];
        const bool first = A.f2c[2 * i] == c;
        const double xf = A.xfc[i], yf = A.yfc[i], zf = A.zfc[i];
        const int b = A.fno[i], n = A.fno[i + 1] - b;
        const int* const fn = A.f2n + b;
        const int p1 = fn[0];
        for (int j = 0; j < n; ++j) {
            const int p2 = fn[j];
            const int p3 = (j == n - 1) ? p1 : fn[j + 1];
            const double x21 = A.x[p2] - xf, y21 = A.y[p2] - yf, z21 = A.z[p2] - zf;
            const double x31 = A.x[p3] - xf, y31 = A.y[p3] - yf, z31 = A.z[p3] - zf;
            const double nx = y21 * z31 - y31 * z21;
            const double ny = z21 * x31 - z31 * x21;
            const double nz = x21 * y31 - x31 * y21;
            const double x41 = xc - xf, y41 = yc - yf, z41 = zc - zf;
            const double xt = A.x[p2] + A.x[p3] + xf + xc;
            const double yt = A.y[p2] + A.y[p3] + yf + yc;
            const double zt = A.z[p2] + A.z[p3] + zf + zc;
            double volt = nx * x41 + ny * y41 + nz * z41;
            if (first) volt = -volt;
            xs += volt * xt;
            ys += volt * yt;
            zs += volt * zt;
            v += volt;
        }
    }
    const double tmp = 1.0 / (4.0 * v + kMetricTiny);
    A.xcct[c] = xs * tmp;
    A.ycct[c] = ys * tmp;
    A.zcct[c] = zs * tmp;
    v /= 6.0;
    A.vol[c] = v;
    if (v < kMetricTiny) atomicAdd(A.small, 1);
}

// CellVolCentroid (:536-571): the volume centroid replaces the averaged
// centre unless it lies on the wrong side of one of the cell's faces
__global__ __launch_bounds__(kBlock) void k_cell_mark(MetricArgs A) {
    const int c = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (c >= A.nC) return;
    const double xt = A.xcct[c], yt = A.ycct[c], zt = A.zcct[c];
    bool keep = true;
    for (int k = A.cfo[c]; k < A.cfo[c + 1]; ++k) {
        const int i = A.cf[k];
        const double x21 = xt - A.xfc[i], y21 = yt - A.yfc[i], z21 = zt - A.zfc[i];
        const double dot = x21 * A.xfn[i] + y21 * A.yfn[i] + z21 * A.zfn[i];
        if (A.f2c[2 * i] == c ? (dot > 0.0) : (dot < 0.0)) keep = false;
    }
    if (keep) {
        A.xcc[c] = xt;
        A.ycc[c] = yt;
        A.zcc[c] = zt;
    }
}

// CellVolCentroid (:572-612): ghost centres mirrored through the boundary
// face, ghost volumes copied
__global__ __launch_bounds__(kBlock) void k_ghost_geom(MetricArgs A) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= A.nB) return;
    const int c1 = A.f2c[2 * i], c2 = A.nC + i;
    if (A.area[i] > kMetricTiny) {
        const double tmp = 2. * ((A.xcc[c1] - A.xfc[i]) * A.xfn[i] + (A.ycc[c1] - A.yfc[i]) * A.yfn[i] +
                                 (A.zcc[c1] - A.zfc[i]) * A.zfn[i]);
        A.xcc[c2] = A.xcc[c1] - A.xfn[i] * tmp;
        A.ycc[c2] = A.ycc[c1] - A.yfn[i] * tmp;
        A.zcc[c2] = A.zcc[c1] - A.zfn[i] * tmp;
    } else {
        A.xcc[c2] = -A.xcc[c1] + 2. * A.xfc[i];
        A.ycc[c2] = -A.ycc[c1] + 2. * A.yfc[i];
        A.zcc[c2] = -A.zcc[c1] + 2. * A.zfc[i];
    }
    A.vol[c2] = A.vol[c1];
}

// CorrectFaceNormal (:619-671): a degenerate face's normal from its first
// cell's centre, made orthogonal to the face's first node
__global__ __launch_bounds__(kBlock) void k_face_normal_fix(MetricArgs A) {
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= A.nF || !(A.area[i] < kMetricTiny)) return;
    const int p1 = A.f2n[A.fno[i]];
    const int c1 = A.f2c[2 * i];
    double x21 = A.xcc[c1] - A.xfc[i], y21 = A.ycc[c1] - A.yfc[i], z21 = A.zcc[c1] - A.zfc[i];
    const double x31 = A.x[p1] - A.xfc[i], y31 = A.y[p1] - A.yfc[i], z31 = A.z[p1] - A.zfc[i];
    const double dot = x21 * x31 + y21 * y31 + z21 * z31;
    double nn = x31 * x31 + y31 * y31 + z31 * z31;
    double ratio = 0.0;
    if (nn > kMetricTiny) ratio = dot / nn;
    x21 -= ratio * x31;
    y21 -= ratio * y31;
    z21 -= ratio * z31;
    nn = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
    A.xfn[i] = -x21 / nn;
    A.yfn[i] = -y21 / nn;
    A.zfn[i] = -z21 / nn;
}

// CorrectCellCentroid (:679-823): a centre closer than 1e-10 to one of its
// faces' planes is pushed off it, face by face in ascending face order
__device__ __forceinline__ double absmax_signfirst(double a, double b) {
    const int sign = (a > 0) ? 1 : -1;
    double c = fmax(fabs(a), fabs(b));
    c *= sign;
    return c;
}

__global__ __launch_bounds__(kBlock) void k_cell_centroid_fix(MetricArgs A, double scale) {
    const int c = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (c >= A.nC) return;
    constexpr double minDn = 1.0e-10;
    double xc = A.xcc[c], yc = A.ycc[c], zc = A.zcc[c];
    for (int k = A.cfo[c]; k < A.cfo[c + 1]; ++k) {
        const int i = A.cf[k];
        const double x21 = xc - A.xfc[i], y21 = yc - A.yfc[i], z21 = zc - A.zfc[i];
        const double dot = x21 * A.xfn[i] + y21 * A.yfn[i] + z21 * A.zfn[i];
        if (fabs(dot) < minDn) {
            double dx = minDn * A.xfn[i], dy = minDn * A.yfn[i], dz = minDn * A.zfn[i];
            dx = absmax_signfirst(dx, scale * xc);
            dy = absmax_signfirst(dy, scale * yc);
            dz = absmax_signfirst(dz, scale * zc);
            const int sign = (dot > 0.0) ? 1 : -1;
            xc += sign * dx;
            yc += sign * dy;
            zc += sign * dz;
        }
    }
    A.xcc[c] = xc;
    A.ycc[c] = yc;
    A.zcc[c] = zc;
}

// The per-slot tables of the residual kernel, the sweeps and the gradient,
// from each cell's slot face ids (its faces in ascending mesh face id):
// the neighbour across the face (the cell itself for an empty slot), the
// face's unit normal pointing away from the cell with its area (negated
// exactly where the cell is the face's second cell; zeros for an empty
// slot), the distance on this cell's side, per cell the slots where it is
// the second cell (bits 0-7) and its face count (bits 8-11), and for
// tetrahedra each slot's three face nodes in the face's order.
template <int LPC>
__global__ __launch_bounds__(kBlock) void k_slot_tables(int nC, const int* __restrict__ cfs,
                                                       const int* __restrict__ cfo, const int2* __restrict__ fcell,
                                                       const double4* __restrict__ fgeo,
                                                       const double2* __restrict__ fdn, const int* __restrict__ fno,
                                                       const int* __restrict__ fn, int* snb, double4* sgeo,
                                                       float4* sgeo32, double* sdn, int* sflip, int* snode) {
    const int d = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (d >= nC) return;
    int flips = (cfo[d + 1] - cfo[d]) << 8;
    for (int sl = 0; sl < LPC; ++sl) {
        const size_t i = (size_t)d * LPC + sl;
        const int f = cfs[i];
        if (f < 0) {
            snb[i] = d;
            sgeo[i] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (sgeo32) sgeo32[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            sdn[i] = 0.0;
            continue;
        }
        const int2 cc = fcell[f];
        const bool flip = cc.y == d;
        if (flip) flips |= 1 << sl;
        const double4 g = fgeo[f];
        const double4 o = flip ? make_double4(-g.x, -g.y, -g.z, g.w) : g;
        snb[i] = flip ? cc.x : cc.y;
        sgeo[i] = o;
        if (sgeo32) sgeo32[i] = make_float4((float)o.x, (float)o.y, (float)o.z, (float)o.w);
        const double2 dn = fdn[f];
        sdn[i] = flip ? dn.y : dn.x;
        if (snode)
            for (int k = 0; k < 3; ++k) snode[12 * (size_t)d + 3 * sl + k] = fn[fno[f] + k];
    }
    sflip[d] = flips;
}
