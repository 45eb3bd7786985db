/*
 * Test infrastructure (oracle): plain-C restatement of the MiniFS solver
 * path.  See mfs_oracle.h for scope.  Every function cites the reference
 * lines it follows; expression trees are kept operand for operand and the
 * file is compiled with -ffp-contract=off so it rounds like the reference.
 */
#define _GNU_SOURCE
#include "mfs_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define BIG 1.e40   /* number_type.h:28 */
#define TINY 1.e-40 /* number_type.h:29 */
#define PI_ 3.14159265358979323846
enum { WALL = 3, SYMM = 4, FAR_FIELD = 6, FAR_FIELD_MOD = 61, INTERFACE = 10 };

static char g_err[256];
const char* or_error(void) { return g_err; }
static int fail(const char* m) {
    snprintf(g_err, sizeof g_err, "%s", m);
    return -1;
}

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

typedef struct {
    double mach, alpha, re, T, gam, gascon, cfl_start, cfl_end, cfl_ratio, epsa_r, lhs_omga, ratio_dtmax;
    double ratio_rhop_max, ratio_rhop_min, ratio_p_break; /* para_field_global.cpp:200-202 */
    int cfl_nstep;
    double u_s, v_s, w_s, p_s, rho_s, p_bar, gamm1, p_min, p_max, rho_min, rho_max, p_break, alf_l;
} Par;

struct or_grid {
    int nN, nF, nC, nB;
    double *x, *y, *z;
    int *nNPF, *fo, *f2n, *f2c;
    int npatch, patch_type[64];
    int* btype;
    double *xfn, *yfn, *zfn, *area, *xfc, *yfc, *zfc;
    double *xcc, *ycc, *zcc, *vol;
    double *q[5]; /* rho u v w p, nC+nB */
    double *res, *dt, *dq, *dqdx, *dqdy, *dqdz;
    /* connectivity */
    int *cfo, *cf;       /* CalC2F: faces per cell, ascending */
    int *cno, *cn;       /* C2N: unique nodes per cell */
    int *Nmark;
    double *WeightNode, *wbf, *wcn; /* WeightNodeBFace2C (per f2n entry), WeightNodeC2N (per cn entry) */
    int nsymm, *symm_node;          /* FindNodeSYMM: nodes on SYMM faces */
    double* symm_n;                 /* their mean face normal, 3 per node */
    /* level schedule for the threaded mode */
    int *flev_off, *flev_cells, nflev;
    int iter_done;
    double dt_max, dt_min, cfl_min;
    int ghost_refresh; /* SetGhostVariables after each update (the commented main.cpp:283 call) */
    double lim[4];     /* ratio_rhop_min, ratio_rhop_max, ratio_p_break, ratio_dtmax (or_set_limits) */
    long long cnt[OR_NCOUNT]; /* how often each guard branch was taken (or_counts) */
    Par P;
};

/* counts a guard branch; the sweeps and the update run in parallel loops in
 * the threaded mode */
static void bump(or_grid* g, int k) {
#pragma omp atomic
    g->cnt[k]++;
}

/* ---------------------------------------------------------------- reader */
/* ReadMFlowGrid_binary (load_grid.cpp:129-273) + ReadMFlowBoundaryCondition (:282-314) */
int or_load(const char* mesh_path, const char* bc_path, or_grid** out) {
    FILE* fp = fopen(mesh_path, "rb");
    if (!fp) return fail("cannot open mesh");
    or_grid* g = calloc(1, sizeof(or_grid));
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, fp) != 3) return fail("short mesh header");
    g->nN = hdr[0];
    g->nF = hdr[1];
    g->nC = hdr[2];
    g->x = malloc(sizeof(double) * g->nN);
    g->y = malloc(sizeof(double) * g->nN);
    g->z = malloc(sizeof(double) * g->nN);
    if (fread(g->x, 8, g->nN, fp) != (size_t)g->nN || fread(g->y, 8, g->nN, fp) != (size_t)g->nN ||
        fread(g->z, 8, g->nN, fp) != (size_t)g->nN)
        return fail("short coordinates");
    g->nNPF = malloc(sizeof(int) * g->nF);
    if (fread(g->nNPF, 4, g->nF, fp) != (size_t)g->nF) return fail("short nNPF");
    g->fo = malloc(sizeof(int) * (g->nF + 1));
    g->fo[0] = 0;
    for (int i = 0; i < g->nF; ++i) g->fo[i + 1] = g->fo[i] + g->nNPF[i];
    g->f2n = malloc(sizeof(int) * g->fo[g->nF]);
    if (fread(g->f2n, 4, g->fo[g->nF], fp) != (size_t)g->fo[g->nF]) return fail("short f2n");
    g->f2c = malloc(sizeof(int) * 2 * g->nF);
    if (fread(g->f2c, 4, 2 * g->nF, fp) != (size_t)(2 * g->nF)) return fail("short f2c");
    fclose(fp);
    int nB = 0;
    for (int i = 0; i < g->nF; ++i) {
        int* c = g->f2c + 2 * i;
        if (c[0] < 0) {
            /* std::reverse(&f2n[facnod[i]+1], &f2n[facnod[i+1]]) and swap */
            int a = g->fo[i] + 1, b = g->fo[i + 1] - 1;
            while (a < b) {
                int t = g->f2n[a];
                g->f2n[a] = g->f2n[b];
                g->f2n[b] = t;
                ++a;
                --b;
            }
            int t = c[0];
            c[0] = c[1];
            c[1] = t;
            ++nB;
        } else if (c[1] < 0) {
            ++nB;
        }
    }
    g->nB = nB;
    FILE* bf = fopen(bc_path, "r");
    if (!bf) return fail("cannot open bc file");
    if (fscanf(bf, "%d", &g->npatch) != 1 || g->npatch < 1 || g->npatch > 64) return fail("bad bc file");
    for (int i = 0; i < g->npatch; ++i) {
        int pid, type;
        if (fscanf(bf, "%d %d", &pid, &type) != 2) return fail("bad bc record");
        g->patch_type[i] = type;
    }
    fclose(bf);
    g->lim[0] = 1.0e-4; /* para_field_global.cpp:200-202, 115 */
    g->lim[1] = 10.0;
    g->lim[2] = 0.01;
    g->lim[3] = 1.0e80;
    *out = g;
    return 0;
}

/* ---------------------------------------------------------------- set-up */
static void set_par(Par* P, double mach, double alpha, double cfl, const double lim[4]) {
    /* para_field_global.cpp:37-209; lim = ratio_rhop_min, ratio_rhop_max,
     * ratio_p_break, ratio_dtmax (the reference's defaults unless set) */
    P->mach = mach;
    P->alpha = alpha;
    P->re = 1.8140449177e7;
    P->T = 288.15;
    P->gam = 1.4;
    P->gascon = 287.053;
    P->cfl_start = cfl;
    P->cfl_end = cfl;
    P->cfl_nstep = 2000;
    P->cfl_ratio = P->cfl_end / P->cfl_start;
    P->epsa_r = 0.1;
    P->alf_l = P->epsa_r;
    P->lhs_omga = 1.0;
    P->ratio_rhop_min = lim ? lim[0] : 1.0e-4;
    P->ratio_rhop_max = lim ? lim[1] : 10.0;
    P->ratio_p_break = lim ? lim[2] : 0.01;
    P->ratio_dtmax = lim ? lim[3] : 1.0e80;
    const double tref = 288.15, sref = 110.4, amuref = 1.78938e-5;
    double trat = (P->T / tref);
    double amu = amuref * trat * sqrt(trat) * (tref + sref) / (P->T + sref);
    double ainf = sqrt(P->gam * P->gascon * P->T);
    double uqq = ainf * P->mach;
    double alpha_t = (P->alpha * PI_ / 180.), beita_t = (0.0 * PI_ / 180.);
    P->u_s = uqq * cos(alpha_t) * cos(-1. * beita_t);
    P->v_s = uqq * sin(-1. * beita_t);
    P->w_s = uqq * sin(alpha_t) * cos(-1. * beita_t);
    P->p_s = 0.00;
    P->rho_s = P->re * amu / uqq;
    P->p_bar = P->rho_s * P->gascon * P->T;
    P->gamm1 = P->gam - 1.0;
    double temp = 1.0 + 0.5 * P->gamm1 * P->mach * P->mach;
    double p_stag_t = pow(temp, P->gam / P->gamm1);
    double rho_stag = pow(temp, 1.0 / P->gamm1) * P->rho_s;
    double p_stag = (p_stag_t - 1.0) * P->p_bar;
    P->p_min = P->ratio_rhop_min * P->p_bar - P->p_bar;
    P->p_max = P->ratio_rhop_max * (p_stag + P->p_bar) - P->p_bar;
    P->rho_min = P->ratio_rhop_min * P->rho_s;
    P->rho_max = P->ratio_rhop_max * rho_stag;
    P->p_break = P->ratio_p_break * P->p_bar - P->p_bar;
}

static int metrics(or_grid* g) {
    /* ComputeMetrics: FaceCellCenterbyAverage (:271-355), FaceAreaNormalCentroid_cycle
     * (:68-266), CellVolCentroid (:357-612), CorrectFaceNormal (:619-671),
     * CorrectCellCentroid (:679-808) of grid_polyhedra_auxiliary_functions.cpp */
    const int nC = g->nC, nB = g->nB, nF = g->nF, n = nC + nB;
    int *f2c = g->f2c, *f2n = g->f2n, *nNPF = g->nNPF;
    double *x = g->x, *y = g->y, *z = g->z;
    double *xfc = g->xfc = calloc(nF, 8), *yfc = g->yfc = calloc(nF, 8), *zfc = g->zfc = calloc(nF, 8);
    double *xfn = g->xfn = calloc(nF, 8), *yfn = g->yfn = calloc(nF, 8), *zfn = g->zfn = calloc(nF, 8);
    double* area = g->area = calloc(nF, 8);
    double *xcc = g->xcc = calloc(n, 8), *ycc = g->ycc = calloc(n, 8), *zcc = g->zcc = calloc(n, 8);
    double* vol = g->vol = calloc(n, 8);
    int node = 0;
    for (int i = 0; i < nF; i++) {
        for (int j = 0; j < nNPF[i]; j++) {
            int p1 = f2n[node++];
            xfc[i] += x[p1];
            yfc[i] += y[p1];
            zfc[i] += z[p1];
        }
        xfc[i] /= nNPF[i];
        yfc[i] /= nNPF[i];
        zfc[i] /= nNPF[i];
    }
    int* nface = calloc(nC, sizeof(int));
    for (int i = 0; i < nF; i++) {
        int c1 = f2c[2 * i];
        xcc[c1] += xfc[i];
        ycc[c1] += yfc[i];
        zcc[c1] += zfc[i];
        nface[c1]++;
        if (i >= nB) {
            int c2 = f2c[2 * i + 1];
            xcc[c2] += xfc[i];
            ycc[c2] += yfc[i];
            zcc[c2] += zfc[i];
            nface[c2]++;
        }
    }
    for (int i = 0; i < nC; i++) {
        xcc[i] /= nface[i];
        ycc[i] /= nface[i];
        zcc[i] /= nface[i];
    }
    free(nface);
    double *xfct = malloc(8 * nF), *yfct = malloc(8 * nF), *zfct = malloc(8 * nF);
    for (int pass = 0; pass <= 4; pass++) {
        int count = 0;
        const int last = (pass == 4);
        for (int i = 0; i < nF; i++) {
            area[i] = 0.0;
            xfct[i] = yfct[i] = zfct[i] = 0.0;
            if (last) xfn[i] = yfn[i] = zfn[i] = 0.0;
            int p1 = f2n[count];
            for (int j = 0; j < nNPF[i]; j++) {
                int p2 = f2n[count++];
                int p3 = (j == nNPF[i] - 1) ? p1 : f2n[count];
                double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
                double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
                double xn = y21 * z31 - y31 * z21, yn = z21 * x31 - z31 * x21, zn = x21 * y31 - x31 * y21;
                double art = sqrt(xn * xn + yn * yn + zn * zn);
                double xt = x[p2] + x[p3] + xfc[i], yt = y[p2] + y[p3] + yfc[i], zt = z[p2] + z[p3] + zfc[i];
                area[i] += art;
                xfct[i] += art * xt;
                yfct[i] += art * yt;
                zfct[i] += art * zt;
                if (last) {
                    xfn[i] += xn;
                    yfn[i] += yn;
                    zfn[i] += zn;
                }
            }
            if (area[i] > TINY) {
                double art = 1.0 / area[i] / 3.0;
                if (!last) {
                    xfct[i] *= art;
                    yfct[i] *= art;
                    zfct[i] *= art;
                    xfc[i] = xfct[i];
                    yfc[i] = yfct[i];
                    zfc[i] = zfct[i];
                } else {
                    xfc[i] = xfct[i] * art;
                    yfc[i] = yfct[i] * art;
                    zfc[i] = zfct[i] * art;
                }
            }
            if (last) {
                double art = sqrt(xfn[i] * xfn[i] + yfn[i] * yfn[i] + zfn[i] * zfn[i]);
                area[i] = art * 0.5;
                if (art > TINY) {
                    xfn[i] /= art;
                    yfn[i] /= art;
                    zfn[i] /= art;
                } else {
                    int c1 = f2c[i + i];
                    double xn = xfc[i] - xcc[c1], yn = yfc[i] - ycc[c1], zn = zfc[i] - zcc[c1];
                    art = sqrt(xn * xn + yn * yn + zn * zn);
                    xfn[i] = xn / art;
                    yfn[i] = yn / art;
                    zfn[i] = zn / art;
                }
            }
        }
    }
    free(xfct);
    free(yfct);
    free(zfct);
    double *xcct = calloc(nC, 8), *ycct = calloc(nC, 8), *zcct = calloc(nC, 8);
    node = 0;
    for (int i = 0; i < nF; i++) {
        int c1 = f2c[2 * i], c2 = f2c[2 * i + 1];
        int p1 = f2n[node];
        for (int j = 0; j < nNPF[i]; j++) {
            int p2 = f2n[node++];
            int p3 = (j == nNPF[i] - 1) ? p1 : f2n[node];
            double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
            double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
            double nx = y21 * z31 - y31 * z21, ny = z21 * x31 - z31 * x21, nz = x21 * y31 - x31 * y21;
            double x41 = xcc[c1] - xfc[i], y41 = ycc[c1] - yfc[i], z41 = zcc[c1] - zfc[i];
            double xt = x[p2] + x[p3] + xfc[i] + xcc[c1], yt = y[p2] + y[p3] + yfc[i] + ycc[c1],
                   zt = z[p2] + z[p3] + zfc[i] + zcc[c1];
            double volt = nx * x41 + ny * y41 + nz * z41;
            volt = -volt;
            xcct[c1] += volt * xt;
            ycct[c1] += volt * yt;
            zcct[c1] += volt * zt;
            vol[c1] += volt;
            if (i >= nB) {
                x41 = xcc[c2] - xfc[i];
                y41 = ycc[c2] - yfc[i];
                z41 = zcc[c2] - zfc[i];
                xt = x[p2] + x[p3] + xfc[i] + xcc[c2];
                yt = y[p2] + y[p3] + yfc[i] + ycc[c2];
                zt = z[p2] + z[p3] + zfc[i] + zcc[c2];
                volt = nx * x41 + ny * y41 + nz * z41;
                xcct[c2] += volt * xt;
                ycct[c2] += volt * yt;
                zcct[c2] += volt * zt;
                vol[c2] += volt;
            }
        }
    }
    int small = 0;
    for (int i = 0; i < nC; i++) {
        double tmp = 1.0 / (4.0 * vol[i] + TINY);
        xcct[i] *= tmp;
        ycct[i] *= tmp;
        zcct[i] *= tmp;
        vol[i] /= 6.0;
        if (vol[i] < TINY) small++;
    }
    if (small) return fail("cell volume below TINY");
    char* mark = malloc(nC);
    memset(mark, 1, nC);
    for (int i = 0; i < nF; i++) {
        int c1 = f2c[i + i];
        double dot = (xcct[c1] - xfc[i]) * xfn[i] + (ycct[c1] - yfc[i]) * yfn[i] + (zcct[c1] - zfc[i]) * zfn[i];
        if (dot > 0.0) mark[c1] = 0;
        if (i >= nB) {
            int c2 = f2c[i + i + 1];
            dot = (xcct[c2] - xfc[i]) * xfn[i] + (ycct[c2] - yfc[i]) * yfn[i] + (zcct[c2] - zfc[i]) * zfn[i];
            if (dot < 0.0) mark[c2] = 0;
        }
    }
    for (int i = 0; i < nC; i++)
        if (mark[i] == 1) {
            xcc[i] = xcct[i];
            ycc[i] = ycct[i];
            zcc[i] = zcct[i];
        }
    free(mark);
    free(xcct);
    free(ycct);
    free(zcct);
    for (int i = 0; i < nB; i++) {
        int c1 = f2c[2 * i], c2 = i + nC;
        if (area[i] > TINY) {
            double tmp = 2. * ((xcc[c1] - xfc[i]) * xfn[i] + (ycc[c1] - yfc[i]) * yfn[i] + (zcc[c1] - zfc[i]) * zfn[i]);
            xcc[c2] = xcc[c1] - xfn[i] * tmp;
            ycc[c2] = ycc[c1] - yfn[i] * tmp;
            zcc[c2] = zcc[c1] - zfn[i] * tmp;
        } else {
            xcc[c2] = -xcc[c1] + 2. * xfc[i];
            ycc[c2] = -ycc[c1] + 2. * yfc[i];
            zcc[c2] = -zcc[c1] + 2. * zfc[i];
        }
        vol[c2] = vol[c1];
    }
    node = 0;
    for (int i = 0; i < nF; i++) { /* CorrectFaceNormal */
        int p1 = f2n[node];
        node += nNPF[i];
        if (area[i] < TINY) {
            int c1 = f2c[i + i];
            double x21 = xcc[c1] - xfc[i], y21 = ycc[c1] - yfc[i], z21 = zcc[c1] - zfc[i];
            double x31 = x[p1] - xfc[i], y31 = y[p1] - yfc[i], z31 = z[p1] - zfc[i];
            double dot = x21 * x31 + y21 * y31 + z21 * z31, nn = x31 * x31 + y31 * y31 + z31 * z31, ratio = 0.0;
            if (nn > TINY) ratio = dot / nn;
            x21 -= ratio * x31;
            y21 -= ratio * y31;
            z21 -= ratio * z31;
            nn = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
            xfn[i] = -x21 / nn;
            yfn[i] = -y21 / nn;
            zfn[i] = -z21 / nn;
        }
    }
    const double scale = pow(10.0, -16 + 1), minDn = 1.0e-10; /* CorrectCellCentroid */
    for (int i = 0; i < nF; i++)
        for (int side = 0; side < (i >= nB ? 2 : 1); ++side) {
            int c = f2c[2 * i + side];
            double dot = (xcc[c] - xfc[i]) * xfn[i] + (ycc[c] - yfc[i]) * yfn[i] + (zcc[c] - zfc[i]) * zfn[i];
            if (fabs(dot) < minDn) {
                double d[3] = {minDn * xfn[i], minDn * yfn[i], minDn * zfn[i]}, cc[3] = {xcc[c], ycc[c], zcc[c]};
                for (int k = 0; k < 3; ++k) { /* AbsMaxSignFirst */
                    int sign = (d[k] > 0) ? 1 : -1;
                    double v = dmax(fabs(d[k]), fabs(scale * cc[k]));
                    d[k] = v * sign;
                }
                int sign = (dot > 0.0) ? 1 : -1;
                xcc[c] += sign * d[0];
                ycc[c] += sign * d[1];
                zcc[c] += sign * d[2];
            }
        }
    return 0;
}

static void ghosts(or_grid* g) {
    /* SetGhostVariables, inviscid (load_grid.cpp:953-1212) */
    const Par* P = &g->P;
    double *rho = g->q[0], *u = g->q[1], *v = g->q[2], *w = g->q[3], *p = g->q[4];
    const double gam = P->gam, p_bar = P->p_bar;
    const double norm_uvw = sqrt(P->u_s * P->u_s + P->v_s * P->v_s + P->w_s * P->w_s);
    for (int i = 0; i < g->nB; i++) {
        int c1 = g->f2c[2 * i], c2 = g->f2c[2 * i + 1], t = g->btype[i];
        double xn = g->xfn[i], yn = g->yfn[i], zn = g->zfn[i];
        if (t == INTERFACE) continue;
        if (t == WALL || t == SYMM) {
            rho[c2] = rho[c1];
            p[c2] = p[c1];
            double vn = 2. * (xn * u[c1] + yn * v[c1] + zn * w[c1]);
            u[c2] = u[c1] - vn * xn;
            v[c2] = v[c1] - vn * yn;
            w[c2] = w[c1] - vn * zn;
        } else if (t == FAR_FIELD) {
            double um = P->u_s, vm = P->v_s, wm = P->w_s, up = u[c1], vp = v[c1], wp = w[c1];
            double rhom = P->rho_s, pm = P->p_s + p_bar, rhop = rho[c1], pp = p[c1] + p_bar;
            double vnm = xn * um + yn * vm + zn * wm, vnp = xn * up + yn * vp + zn * wp;
            double cm = sqrt(gam * pm / rhom), cp = sqrt(gam * pp / rhop);
            double riemm = vnm - 2. * cm / (gam - 1.), riemp = vnp + 2. * cp / (gam - 1.);
            double vnb = 0.5 * (riemp + riemm), cb = 0.25 * (riemp - riemm) * (gam - 1.);
            if (fabs(vnb / cb) > 1.) {
                if (vnb <= 0.0) {
                    rho[c2] = rhom; u[c2] = um; v[c2] = vm; w[c2] = wm; p[c2] = pm;
                } else {
                    rho[c2] = rhop; u[c2] = up; v[c2] = vp; w[c2] = wp; p[c2] = pp;
                }
            } else {
                double rel = vnb / norm_uvw;
                if (rel <= -0.0 || rel > 0.0) {
                    int in = rel <= -0.0;
                    double bu = in ? um : up, bv = in ? vm : vp, bw = in ? wm : wp, bvn = in ? vnm : vnp;
                    double bp = in ? pm : pp, brho = in ? rhom : rhop;
                    double vtx = bu - bvn * xn, vty = bv - bvn * yn, vtz = bw - bvn * zn;
                    double entr = bp / pow(brho, gam);
                    rho[c2] = pow((cb * cb / (entr * gam)), (double)(1. / (gam - 1.)));
                    u[c2] = vtx + vnb * xn;
                    v[c2] = vty + vnb * yn;
                    w[c2] = vtz + vnb * zn;
                    p[c2] = cb * cb * rho[c2] / gam;
                } else {
                    rho[c2] = 0.5 * (rhop + rhom); u[c2] = 0.5 * (up + um); v[c2] = 0.5 * (vp + vm);
                    w[c2] = 0.5 * (wp + wm); p[c2] = 0.5 * (pp + pm);
                }
            }
            rho[c2] = 2 * rho[c2] - rhop;
            u[c2] = 2 * u[c2] - up;
            v[c2] = 2 * v[c2] - vp;
            w[c2] = 2 * w[c2] - wp;
            p[c2] = 2 * p[c2] - pp;
            p[c2] -= p_bar;
        }
        if (rho[c2] < P->rho_min) bump(g, OR_GHOST_RHO_LO);
        if (rho[c2] > P->rho_max) bump(g, OR_GHOST_RHO_HI);
        rho[c2] = dmax(rho[c2], P->rho_min);
        rho[c2] = dmin(rho[c2], P->rho_max);
        if (p[c2] < P->p_min) bump(g, OR_GHOST_P_LO);
        if (p[c2] > P->p_max) bump(g, OR_GHOST_P_HI);
        p[c2] = dmax(p[c2], P->p_min);
        p[c2] = dmin(p[c2], P->p_max);
    }
}

static void connectivity_and_weights(or_grid* g) {
    const int nC = g->nC, nF = g->nF, nB = g->nB, nN = g->nN;
    /* CalnFPC / CalC2F (grid_polyhedra_auxiliary_functions.cpp:1293-1399) */
    g->cfo = calloc(nC + 1, sizeof(int));
    for (int i = 0; i < nF; ++i) {
        g->cfo[g->f2c[2 * i] + 1]++;
        if (i >= nB) g->cfo[g->f2c[2 * i + 1] + 1]++;
    }
    for (int c = 0; c < nC; ++c) g->cfo[c + 1] += g->cfo[c];
    g->cf = malloc(sizeof(int) * g->cfo[nC]);
    int* fill = malloc(sizeof(int) * nC);
    memcpy(fill, g->cfo, sizeof(int) * nC);
    for (int i = 0; i < nF; ++i) {
        g->cf[fill[g->f2c[2 * i]]++] = i;
        if (i >= nB) g->cf[fill[g->f2c[2 * i + 1]]++] = i;
    }
    /* unique cell nodes (C2N), in first-appearance order over the cell's faces */
    g->cno = calloc(nC + 1, sizeof(int));
    int* tmp = malloc(sizeof(int) * 64);
    g->cn = malloc(sizeof(int) * (size_t)g->cfo[nC] * 4 + 1);
    int pos = 0;
    for (int c = 0; c < nC; ++c) {
        int k = 0;
        for (int e = g->cfo[c]; e < g->cfo[c + 1]; ++e) {
            int f = g->cf[e];
            for (int j = g->fo[f]; j < g->fo[f + 1]; ++j) {
                int p1 = g->f2n[j], dup = 0;
                for (int t = 0; t < k; ++t) dup |= tmp[t] == p1;
                if (!dup && k < 64) tmp[k++] = p1;
            }
        }
        for (int t = 0; t < k; ++t) g->cn[pos++] = tmp[t];
        g->cno[c + 1] = pos;
    }
    free(tmp);
    free(fill);
    /* ComputeWeight3D_Node (:831-1049) */
    g->Nmark = calloc(nN, sizeof(int));
    g->WeightNode = calloc(nN, 8);
    g->wbf = calloc(g->fo[nB] + 1, 8);
    g->wcn = calloc(pos + 1, 8);
    for (int pass = 0; pass < 3; ++pass)
        for (int i = 0; i < nB; i++) {
            int t = g->btype[i];
            int take = pass == 0 ? t == WALL
                     : pass == 1 ? (t == FAR_FIELD || t == FAR_FIELD_MOD)
                                 : !(t == WALL || t == SYMM || t == FAR_FIELD || t == FAR_FIELD_MOD || t == INTERFACE);
            if (!take) continue;
            for (int j = g->fo[i]; j < g->fo[i + 1]; j++) {
                int p1 = g->f2n[j];
                if (pass == 1 && g->Nmark[p1] == WALL) continue;
                if (pass == 2 && (g->Nmark[p1] == WALL || g->Nmark[p1] == FAR_FIELD)) continue;
                double dx = g->x[p1] - g->xfc[i], dy = g->y[p1] - g->yfc[i], dz = g->z[p1] - g->zfc[i];
                double wr = dx * dx + dy * dy + dz * dz;
                wr = sqrt(wr);
                wr = 1. / wr;
                g->wbf[j] = wr;
                g->WeightNode[p1] += wr;
                g->Nmark[p1] = pass == 0 ? WALL : pass == 1 ? FAR_FIELD : 101;
            }
        }
    for (int c = 0; c < nC; ++c)
        for (int e = g->cno[c]; e < g->cno[c + 1]; ++e) {
            int p1 = g->cn[e];
            if (g->Nmark[p1] != 0) continue;
            double dx = g->x[p1] - g->xcc[c], dy = g->y[p1] - g->ycc[c], dz = g->z[p1] - g->zcc[c];
            double wr = dx * dx + dy * dy + dz * dz;
            wr = sqrt(wr);
            wr = 1. / wr;
            g->wcn[e] = wr;
            g->WeightNode[p1] += wr;
        }
    /* FindNodeSYMM (grid_polyhedra_auxiliary_functions.cpp:2060-2139):
     * nodes of SYMM faces with the mean of those faces' normals */
    {
        double* sn = calloc(4 * (size_t)nN, 8); /* x, y, z sums and face count */
        int* mark = calloc(nN, sizeof(int));
        for (int i = 0; i < nB; i++) {
            if (g->btype[i] != SYMM) continue;
            for (int j = g->fo[i]; j < g->fo[i + 1]; j++) {
                int p1 = g->f2n[j];
                mark[p1] = 1;
                sn[4 * (size_t)p1] += g->xfn[i];
                sn[4 * (size_t)p1 + 1] += g->yfn[i];
                sn[4 * (size_t)p1 + 2] += g->zfn[i];
                sn[4 * (size_t)p1 + 3] += 1.0;
            }
        }
        g->nsymm = 0;
        for (int i = 0; i < nN; ++i) g->nsymm += mark[i];
        g->symm_node = malloc(sizeof(int) * (g->nsymm + 1));
        g->symm_n = malloc(8 * (3 * (size_t)g->nsymm + 1));
        int k = 0;
        for (int i = 0; i < nN; ++i) {
            if (!mark[i]) continue;
            g->symm_node[k] = i;
            for (int a = 0; a < 3; ++a) g->symm_n[3 * k + a] = sn[4 * (size_t)i + a] / sn[4 * (size_t)i + 3];
            ++k;
        }
        free(mark);
        free(sn);
    }
    /* forward level schedule for the threaded mode */
    int* lev = calloc(nC, sizeof(int));
    int nl = 0;
    for (int c = 0; c < nC; ++c) {
        int l = 0;
        for (int e = g->cfo[c]; e < g->cfo[c + 1]; ++e) {
            int f = g->cf[e];
            if (f < nB) continue;
            int nb = g->f2c[2 * f] == c ? g->f2c[2 * f + 1] : g->f2c[2 * f];
            if (nb < c && lev[nb] + 1 > l) l = lev[nb] + 1;
        }
        lev[c] = l;
        if (l + 1 > nl) nl = l + 1;
    }
    g->nflev = nl;
    g->flev_off = calloc(nl + 1, sizeof(int));
    for (int c = 0; c < nC; ++c) g->flev_off[lev[c] + 1]++;
    for (int l = 0; l < nl; ++l) g->flev_off[l + 1] += g->flev_off[l];
    g->flev_cells = malloc(sizeof(int) * nC);
    int* f2 = malloc(sizeof(int) * nl);
    memcpy(f2, g->flev_off, sizeof(int) * nl);
    for (int c = 0; c < nC; ++c) g->flev_cells[f2[lev[c]]++] = c;
    free(f2);
    free(lev);
}

int or_preprocess(or_grid* g, double mach, double alpha, double cfl) {
    set_par(&g->P, mach, alpha, cfl, g->lim);
    const int nC = g->nC, nB = g->nB, n = nC + nB;
    /* SpecifyBC (load_grid.cpp:857-902) */
    g->btype = malloc(sizeof(int) * (nB + 1));
    for (int k = 0; k < nB; k++) {
        int patch = -g->f2c[2 * k + 1];
        if (patch < 1 || patch > g->npatch) return fail("boundary face names an unknown patch");
        g->btype[k] = g->patch_type[patch - 1];
        g->f2c[2 * k + 1] = k + nC;
    }
    if (metrics(g)) return -1;
    /* AllocAndInitFlowfield (load_grid.cpp:571-601) */
    const double qs[5] = {g->P.rho_s, g->P.u_s, g->P.v_s, g->P.w_s, g->P.p_s};
    for (int j = 0; j < 5; ++j) {
        g->q[j] = malloc(8 * n);
        for (int i = 0; i < n; ++i) g->q[j][i] = qs[j];
    }
    ghosts(g);
    connectivity_and_weights(g);
    g->res = calloc(5 * (size_t)nC, 8);
    g->dt = calloc(nC, 8);
    g->dq = calloc(5 * (size_t)n, 8);
    g->dqdx = calloc(5 * (size_t)n, 8);
    g->dqdy = calloc(5 * (size_t)n, 8);
    g->dqdz = calloc(5 * (size_t)n, 8);
    g->iter_done = 0;
    g->dt_max = 0.0;
    g->dt_min = BIG;
    return 0;
}

/* ---------------------------------------------------------------- kernels */
/* RoeFlux_noprec live loop (cal_flux.cpp:682-833) for one face */
static void roe(const double ql[5], const double qr[5], double ax, double ay, double az, double area, const Par* P,
                double f[5]) {
    const double gamm1 = P->gamm1, p_bar = P->p_bar;
    double et_l = (ql[4] + p_bar) / gamm1 + 0.5 * ql[0] * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
    double et_r = (qr[4] + p_bar) / gamm1 + 0.5 * qr[0] * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    double ht_l = et_l + ql[4] + p_bar, ht_r = et_r + qr[4] + p_bar;
    double vn_l = ax * ql[1] + ay * ql[2] + az * ql[3], vn_r = ax * qr[1] + ay * qr[2] + az * qr[3];
    double tmp0 = vn_l * ql[0], tmp1 = vn_r * qr[0];
    f[0] = tmp0 + tmp1;
    f[1] = tmp0 * ql[1] + ax * ql[4] + tmp1 * qr[1] + ax * qr[4];
    f[2] = tmp0 * ql[2] + ay * ql[4] + tmp1 * qr[2] + ay * qr[4];
    f[3] = tmp0 * ql[3] + az * ql[4] + tmp1 * qr[3] + az * qr[4];
    f[4] = ht_l * vn_l + ht_r * vn_r;
    tmp0 = sqrt(qr[0] / ql[0]);
    tmp1 = 1.0 / (1.0 + tmp0);
    double rho_a = sqrt(qr[0] * ql[0]);
    double u_a = (ql[1] + qr[1] * tmp0) * tmp1, v_a = (ql[2] + qr[2] * tmp0) * tmp1, w_a = (ql[3] + qr[3] * tmp0) * tmp1;
    double vn_a = u_a * ax + v_a * ay + w_a * az;
    double h_a = (ht_l / ql[0] + ht_r / qr[0] * tmp0) * tmp1;
    double q2 = 0.5 * (u_a * u_a + v_a * v_a + w_a * w_a);
    double c2_a = gamm1 * (h_a - q2);
    c2_a = fabs(c2_a);
    double c_a = sqrt(c2_a);
    double e1 = fabs(vn_a), e2 = fabs(vn_a + c_a), e3 = fabs(vn_a - c_a);
    double spectral = fabs(u_a) + fabs(v_a) + fabs(w_a) + c_a;
    double epsaa = P->alf_l * spectral, epsbb = 0.25 / dmax(epsaa, TINY), epscc = 2.0 * epsaa;
    if (e1 < epscc) e1 = e1 * e1 * epsbb + epsaa;
    if (e2 < epscc) e2 = e2 * e2 * epsbb + epsaa;
    if (e3 < epscc) e3 = e3 * e3 * epsbb + epsaa;
    double drho = qr[0] - ql[0], du = qr[1] - ql[1], dv = qr[2] - ql[2], dw = qr[3] - ql[3], dp = qr[4] - ql[4];
    double dvn = vn_r - vn_l, dq2 = u_a * du + v_a * dv + w_a * dw;
    tmp0 = dp / c2_a;
    tmp1 = rho_a * dvn / c_a;
    double a1 = (drho - tmp0) * e1, a2 = 0.5 * (tmp0 + tmp1) * e2, a3 = 0.5 * (tmp0 - tmp1) * e3;
    tmp0 = a1 + a2 + a3;
    tmp1 = e1 * rho_a;
    double tmp2 = -tmp1 * dvn + (a2 - a3) * c_a;
    f[0] -= tmp0;
    f[1] -= tmp0 * u_a + tmp1 * du + tmp2 * ax;
    f[2] -= tmp0 * v_a + tmp1 * dv + tmp2 * ay;
    f[3] -= tmp0 * w_a + tmp1 * dw + tmp2 * az;
    f[4] -= a1 * q2 + (a2 + a3) * h_a + tmp1 * dq2 + tmp2 * vn_a;
    tmp0 = 0.5 * area;
    for (int j = 0; j < 5; ++j) f[j] *= tmp0;
}

/* FluxLUSGS3D (forward_lusgs.cpp:1271-1324) */
static void flux_lusgs(const double q[5], const double DQ[5], double nx, double ny, double nz, const Par* P,
                       double flux[5]) {
    double Q[5], gam = P->gam, gam1 = gam - 1., p_bar = P->p_bar;
    Q[0] = q[0];
    Q[1] = q[0] * q[1];
    Q[2] = q[0] * q[2];
    Q[3] = q[0] * q[3];
    double rv2 = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    double p = q[4];
    Q[4] = p / gam1 + rv2;
    double v_n = q[1] * nx + q[2] * ny + q[3] * nz;
    double c2 = gam * (p + p_bar) / q[0];
    double eig = fabs(v_n) + sqrt(c2);
    eig *= P->lhs_omga;
    double peff = gam * p_bar / gam1;
    flux[1] = -Q[1] * v_n - p * nx;
    flux[2] = -Q[2] * v_n - p * ny;
    flux[3] = -Q[3] * v_n - p * nz;
    flux[4] = -(Q[4] + p + peff) * v_n;
    for (int i = 0; i < 5; i++) Q[i] += DQ[i];
    rv2 = 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) / Q[0];
    p = gam1 * (Q[4] - rv2);
    flux[0] = DQ[1] * nx + DQ[2] * ny + DQ[3] * nz;
    v_n *= q[0];
    v_n += flux[0];
    v_n /= Q[0];
    flux[1] += Q[1] * v_n + p * nx;
    flux[2] += Q[2] * v_n + p * ny;
    flux[3] += Q[3] * v_n + p * nz;
    flux[4] += (Q[4] + p + peff) * v_n;
    for (int i = 0; i < 5; i++) flux[i] -= eig * DQ[i];
}

/* DQ_limit == 2 (forward_lusgs.cpp:939-973, 1184-1218) */
static void limiter(or_grid* g, int cell, int fwd) {
    const Par* P = &g->P;
    const size_t n = (size_t)g->nC + g->nB;
    double* const* q = g->q;
    double* DQ = g->dq;
    double vv = q[1][cell] * q[1][cell] + q[2][cell] * q[2][cell] + q[3][cell] * q[3][cell];
    for (int stage = 0; stage < 3; ++stage) {
        double dp = DQ[4 * n + cell] + 0.5 * DQ[cell] * vv -
                    (DQ[n + cell] * q[1][cell] + DQ[2 * n + cell] * q[2][cell] + DQ[3 * n + cell] * q[3][cell]);
        dp *= P->gamm1;
        if ((q[0][cell] + DQ[cell]) < P->rho_min || (q[0][cell] + DQ[cell]) > P->rho_max ||
            (q[4][cell] + dp) < P->p_min || (q[4][cell] + dp) > P->p_max) {
            bump(g, (fwd ? OR_FWD_LIMIT1 : OR_BWD_LIMIT1) + stage);
            for (int i = 0; i < 5; ++i) DQ[i * n + cell] = stage < 2 ? DQ[i * n + cell] * 0.1 : 0.0;
        }
    }
}

/* one cell of the forward (fwd=1) or backward sweep (forward_lusgs.cpp:825-1014, 1060-1249) */
static void sweep_cell(or_grid* g, const double* Diag, int cell, int fwd) {
    const Par* P = &g->P;
    const size_t n = (size_t)g->nC + g->nB;
    double flux_s[5] = {0, 0, 0, 0, 0};
    for (int e = g->cfo[cell]; e < g->cfo[cell + 1]; ++e) {
        int face = g->cf[e], c1 = g->f2c[2 * face], c2 = g->f2c[2 * face + 1];
        if (fwd ? (c1 > cell || c2 > cell) : (c1 < cell || c2 < cell)) continue;
        double nrm[3] = {g->xfn[face], g->yfn[face], g->zfn[face]};
        if (c2 == cell) {
            c2 = c1;
            nrm[0] = -nrm[0];
            nrm[1] = -nrm[1];
            nrm[2] = -nrm[2];
        }
        double ql[5], dql[5], fl[5];
        for (int i = 0; i < 5; ++i) {
            ql[i] = g->q[i][c2];
            dql[i] = g->dq[i * n + c2];
        }
        flux_lusgs(ql, dql, nrm[0], nrm[1], nrm[2], P, fl);
        if (fwd) {
            double tmp = 0.5 * g->area[face];
            for (int i = 0; i < 5; i++) g->dq[i * n + cell] -= tmp * fl[i];
        } else {
            double tmp = g->area[face];
            for (int i = 0; i < 5; i++) flux_s[i] += tmp * fl[i];
        }
    }
    if (fwd) {
        for (int i = 0; i < 5; i++) g->dq[i * n + cell] /= Diag[cell];
    } else {
        double tmp = 2.0 * Diag[cell];
        for (int i = 0; i < 5; i++) g->dq[i * n + cell] -= flux_s[i] / tmp;
    }
    limiter(g, cell, fwd);
}

/* CompNodeVar3D_dist (cal_gradient.cpp:1922-2327) up to the division: the
 * distance-weighted node values of one variable from the boundary faces
 * (WALL, FAR_FIELD, other types, :2155-2240) and the cells (:2260-2280). */
static void node_values(const or_grid* g, const double* qv, double* qn) {
    const int nN = g->nN, nB = g->nB, nC = g->nC;
    memset(qn, 0, 8 * (size_t)nN);
    for (int pass = 0; pass < 3; ++pass)
        for (int i = 0; i < nB; i++) {
            int t = g->btype[i];
            int take = pass == 0 ? t == WALL
                     : pass == 1 ? t == FAR_FIELD
                                 : !(t == WALL || t == SYMM || t == FAR_FIELD || t == INTERFACE);
            if (!take) continue;
            double fq = qv[g->f2c[2 * i]] + qv[g->f2c[2 * i + 1]];
            fq *= 0.5;
            for (int j = g->fo[i]; j < g->fo[i + 1]; j++) {
                int p1 = g->f2n[j];
                if (pass == 1 && g->Nmark[p1] == WALL) continue;
                if (pass == 2 && (g->Nmark[p1] == WALL || g->Nmark[p1] == FAR_FIELD)) continue;
                qn[p1] += fq * g->wbf[j];
            }
        }
    for (int c = 0; c < nC; ++c)
        for (int e = g->cno[c]; e < g->cno[c + 1]; ++e) {
            int p1 = g->cn[e];
            if (g->Nmark[p1] != 0) continue;
            qn[p1] += qv[c] * g->wcn[e];
        }
    for (int i = 0; i < nN; ++i) qn[i] /= (g->WeightNode[i] + TINY);
}

static void run_iteration(or_grid* g, int threads) {
    const Par* P = &g->P;
    const int nC = g->nC, nB = g->nB, nF = g->nF, nN = g->nN;
    const size_t n = (size_t)nC + nB;
    double* const* q = g->q;
    const int par = threads > 0;
    /* ZeroResiduals + InviscidFlux (cal_flux.cpp:54-192, 1219-1254) */
    if (!par) {
        memset(g->res, 0, sizeof(double) * 5 * nC);
        for (int i = 0; i < nF; ++i) {
            int c1 = g->f2c[2 * i], c2 = g->f2c[2 * i + 1];
            double ql[5], qr[5], F[5];
            for (int j = 0; j < 5; ++j) {
                ql[j] = q[j][c1];
                qr[j] = q[j][c2];
            }
            roe(ql, qr, g->xfn[i], g->yfn[i], g->zfn[i], g->area[i], P, F);
            for (int j = 0; j < 5; ++j) {
                g->res[(size_t)j * nC + c1] -= F[j];
                if (i >= nB) g->res[(size_t)j * nC + c2] += F[j];
            }
        }
    } else {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int c = 0; c < nC; ++c) {
            double r[5] = {0, 0, 0, 0, 0};
            for (int e = g->cfo[c]; e < g->cfo[c + 1]; ++e) {
                int i = g->cf[e], c1 = g->f2c[2 * i], c2 = g->f2c[2 * i + 1];
                double ql[5], qr[5], F[5];
                for (int j = 0; j < 5; ++j) {
                    ql[j] = q[j][c1];
                    qr[j] = q[j][c2];
                }
                roe(ql, qr, g->xfn[i], g->yfn[i], g->zfn[i], g->area[i], P, F);
                for (int j = 0; j < 5; ++j) r[j] = (c1 == c) ? r[j] - F[j] : r[j] + F[j];
            }
            for (int j = 0; j < 5; ++j) g->res[(size_t)j * nC + c] = r[j];
        }
    }
    /* ComputeTimeStep: TimeStepNormal_new, CellIsMG, LimitTimeStep (compute_timestep.cpp:66-998, 1315-1338) */
    double* dt = g->dt;
    int* det = malloc(sizeof(int) * nC);
#pragma omp parallel for num_threads(par ? threads : 1) schedule(static)
    for (int c = 0; c < nC; ++c) {
        double d = BIG;
        int dd = 1;
        for (int e = g->cfo[c]; e < g->cfo[c + 1]; ++e) {
            int i = g->cf[e], c1 = g->f2c[2 * i], c2 = g->f2c[2 * i + 1];
            double c2tmp = P->gam * (q[4][c] + P->p_bar) / q[0][c];
            double dn = fabs((g->xfc[i] - g->xcc[c]) * g->xfn[i] + (g->yfc[i] - g->ycc[c]) * g->yfn[i] +
                             (g->zfc[i] - g->zcc[c]) * g->zfn[i]);
            double vn = q[1][c] * g->xfn[i] + q[2][c] * g->yfn[i] + q[3][c] * g->zfn[i];
            vn = fabs(vn);
            double eigv = vn + sqrt(c2tmp);
            d = dmin(d, dn / eigv);
            double dp = fabs(q[4][c2] - q[4][c1]) / (q[4][c2] + q[4][c1] + P->p_bar + P->p_bar);
            if (dp > 0.1) dd = 0;
        }
        dt[c] = d;
        det[c] = dd;
    }
    double cfl = g->iter_done > P->cfl_nstep ? P->cfl_end
                                             : P->cfl_start * pow(P->cfl_ratio, (double)g->iter_done / P->cfl_nstep);
    g->cfl_min = 0.5 * cfl;
    for (int i = 0; i < nC; ++i) {
        double p = q[4][i], ct;
        if (p > P->p_break)
            ct = cfl;
        else if (p < P->p_min)
            ct = g->cfl_min, bump(g, OR_CFL_PMIN);
        else
            ct = (p - P->p_min) / (P->p_break - P->p_min) * (cfl - g->cfl_min) + g->cfl_min, bump(g, OR_CFL_RAMP);
        if (!det[i]) ct = g->cfl_min, bump(g, OR_CFL_DET);
        dt[i] *= ct;
    }
    free(det);
    for (int i = 0; i < nC; ++i) {
        g->dt_max = dmax(g->dt_max, dt[i]);
        g->dt_min = dmin(g->dt_min, dt[i]);
    }
    if (g->dt_max / g->dt_min > P->ratio_dtmax) {
        double lim = P->ratio_dtmax * g->dt_min;
        for (int i = 0; i < nC; ++i)
            if (dt[i] > lim) dt[i] = lim, bump(g, OR_DTMAX);
    }
    /* ForwardLUSGS (forward_lusgs.cpp:67-163): Diag, CalDiagLUSGS (:554-595) */
    double* Diag = malloc(8 * nC);
#pragma omp parallel for num_threads(par ? threads : 1) schedule(static)
    for (int c = 0; c < nC; ++c) {
        double dg = g->vol[c] / dt[c];
        for (int e = g->cfo[c]; e < g->cfo[c + 1]; ++e) {
            int i = g->cf[e];
            double vn = q[1][c] * g->xfn[i] + q[2][c] * g->yfn[i] + q[3][c] * g->zfn[i];
            vn = fabs(vn);
            double ss = P->gam * (q[4][c] + P->p_bar) / q[0][c];
            double eig = vn + sqrt(ss);
            dg += 0.5 * g->area[i] * eig * P->lhs_omga;
        }
        Diag[c] = dg;
    }
    memset(g->dq, 0, sizeof(double) * 5 * n);
    for (int j = 0; j < 5; ++j) memcpy(g->dq + j * n, g->res + (size_t)j * nC, sizeof(double) * nC);
    /* SolveLUSGS3D (:803-1249) */
    for (int i = 0; i < 5; i++) g->dq[i * n] /= Diag[0];
    if (!par) {
        for (int c = 1; c < nC; ++c) sweep_cell(g, Diag, c, 1);
        for (int c = nC - 1; c >= 0; --c) sweep_cell(g, Diag, c, 0);
    } else {
        for (int l = 0; l < g->nflev; ++l) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int k = g->flev_off[l]; k < g->flev_off[l + 1]; ++k)
                if (g->flev_cells[k] != 0) sweep_cell(g, Diag, g->flev_cells[k], 1);
        }
        /* the backward sweep of cell c needs its higher neighbours, whose
         * forward levels are all larger: descending forward levels are a
         * valid order for it */
        for (int l = g->nflev - 1; l >= 0; --l) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int k = g->flev_off[l]; k < g->flev_off[l + 1]; ++k) sweep_cell(g, Diag, g->flev_cells[k], 0);
        }
    }
    free(Diag);
    /* UpdateFlowField3D_CFL3d (:1335-1488) */
    const double alpq = -0.2, phiq = 1. / 0.5, betq = 1.0 + alpq * phiq, gam1 = P->gam - 1.;
#pragma omp parallel for num_threads(par ? threads : 1) schedule(static)
    for (int i = 0; i < nC; ++i) {
        double *rho = q[0], *u = q[1], *v = q[2], *w = q[3], *p = q[4];
        double rhot = rho[i], ru = rhot * u[i], rv = rhot * v[i], rw = rhot * w[i];
        double re = p[i] / gam1 + 0.5 * rhot * (u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
        rhot += g->dq[i];
        ru += g->dq[n + i];
        rv += g->dq[2 * n + i];
        rw += g->dq[3 * n + i];
        re += g->dq[4 * n + i];
        double rhotr = 1. / (rhot + TINY);
        u[i] = ru * rhotr;
        v[i] = rv * rhotr;
        w[i] = rw * rhotr;
        double rho_del = g->dq[i], rho_rat = rho_del / rho[i];
        if (rho_rat < alpq) rho_del /= betq + fabs(rho_rat) * phiq, bump(g, OR_RELAX_RHO);
        rho[i] += rho_del;
        if (rho[i] < P->rho_min) bump(g, OR_CLAMP_RHO_LO);
        if (rho[i] > P->rho_max) bump(g, OR_CLAMP_RHO_HI);
        rho[i] = dmax(rho[i], P->rho_min);
        rho[i] = dmin(rho[i], P->rho_max);
        double ptmp = gam1 * (re - 0.5 * (u[i] * u[i] + v[i] * v[i] + w[i] * w[i]) * rho[i]);
        double p_del = ptmp - p[i], p_rat = p_del / (p[i] + P->p_bar);
        if (p_rat < alpq) p_del /= betq + fabs(p_rat) * phiq, bump(g, OR_RELAX_P);
        p[i] += p_del;
        if (p[i] < P->p_min) bump(g, OR_CLAMP_P_LO);
        if (p[i] > P->p_max) bump(g, OR_CLAMP_P_HI);
        p[i] = dmax(p[i], P->p_min);
        p[i] = dmin(p[i], P->p_max);
    }
    /* PassInterfaceData: single domain, nothing to pass.  SetGhostVariables
     * when enabled, where main.cpp:283 has it commented out. */
    if (g->ghost_refresh) ghosts(g);
    /* CalculateGradient (cal_gradient.cpp:87-167): CompNodeVar3D_dist + CompGradientQ_Gauss_Node, per variable */
    double* qn = malloc(8 * (size_t)nN);
    memset(g->dqdx, 0, 8 * 5 * n);
    memset(g->dqdy, 0, 8 * 5 * n);
    memset(g->dqdz, 0, 8 * 5 * n);
    double* ru = g->nsymm ? malloc(24 * (size_t)nN) : NULL; /* uncorrected node u, v, w */
    double *rv = ru ? ru + nN : NULL, *rw = ru ? ru + 2 * (size_t)nN : NULL;
    for (int var = 0; var < 5; ++var) {
        const double* qv = q[var];
        node_values(g, qv, qn);
        if (g->nsymm && var >= 1 && var <= 3) {
            /* symmetry-plane nodes (:2332-2348): the node velocity loses its
             * component along the mean SYMM normal.  vn uses the uncorrected
             * node u, v, w, all three formed in the u call (:1989-2325). */
            if (var == 1) {
                memcpy(ru, qn, 8 * (size_t)nN);
                node_values(g, q[2], rv);
                node_values(g, q[3], rw);
            }
            const double* raw = var == 1 ? ru : var == 2 ? rv : rw;
            for (int k = 0; k < g->nsymm; ++k) {
                const int i = g->symm_node[k];
                const double* nrm = g->symm_n + 3 * k;
                double vn = ru[i] * nrm[0] + rv[i] * nrm[1] + rw[i] * nrm[2];
                qn[i] = raw[i] - vn * nrm[var - 1];
            }
        }
        double *gx = g->dqdx + var * n, *gy = g->dqdy + var * n, *gz = g->dqdz + var * n;
        for (int i = 0; i < nF; i++) { /* face loop (:1740-1795) */
            int c1 = g->f2c[2 * i], c2 = g->f2c[2 * i + 1];
            double qsum = 0.0;
            int t = i < nB ? g->btype[i] : -1;
            if (i >= nB || t == INTERFACE || t == SYMM) {
                for (int j = g->fo[i]; j < g->fo[i + 1]; j++) qsum += qn[g->f2n[j]];
                qsum /= (double)g->nNPF[i];
            } else {
                qsum = 0.5 * (qv[c1] + qv[c2]);
            }
            qsum *= g->area[i];
            double tx = qsum * g->xfn[i], ty = qsum * g->yfn[i], tz = qsum * g->zfn[i];
            gx[c1] += tx;
            gy[c1] += ty;
            gz[c1] += tz;
            if (i >= nB) {
                gx[c2] -= tx;
                gy[c2] -= ty;
                gz[c2] -= tz;
            }
        }
        for (int c = 0; c < nC; ++c) {
            gx[c] /= g->vol[c];
            gy[c] /= g->vol[c];
            gz[c] /= g->vol[c];
        }
    }
    free(qn);
    free(ru);
    g->iter_done++;
}

int or_run(or_grid* g, int iters, int threads, double* hist, double* t_loop) {
    if (!g->res) return fail("grid not preprocessed");
    double total = 0.0;
    for (int it = 0; it < iters; ++it) {
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        run_iteration(g, threads);
        clock_gettime(CLOCK_MONOTONIC, &b);
        total += (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
        if (hist) {
            for (int j = 0; j < 5; ++j) { /* main.cpp:294-317 */
                double s = 0.0;
                for (int i = 0; i < g->nC; ++i) s += g->res[(size_t)j * g->nC + i] * g->res[(size_t)j * g->nC + i];
                hist[7 * it + j] = sqrt(s);
            }
            hist[7 * it + 5] = g->dt_max;
            hist[7 * it + 6] = g->dt_min;
        }
    }
    if (t_loop) *t_loop = total;
    return 0;
}

int or_set_ghost_refresh(or_grid* g, int on) {
    g->ghost_refresh = on ? 1 : 0;
    return 0;
}

int or_set(or_grid* g, const char* name, const double* src, size_t count) {
    const char* st[5] = {"rho", "u", "v", "w", "p"};
    const size_t n = (size_t)g->nC + g->nB;
    for (int j = 0; j < 5; ++j)
        if (strcmp(name, st[j]) == 0) {
            if (count != n) return fail("or_set: size mismatch");
            memcpy(g->q[j], src, sizeof(double) * n);
            return 0;
        }
    return fail("or_set: only the state arrays rho u v w p can be set");
}

int or_set_ghosts(or_grid* g) {
    ghosts(g);
    return 0;
}

int or_set_limits(or_grid* g, double ratio_rhop_min, double ratio_rhop_max, double ratio_p_break, double ratio_dtmax) {
    if (!(ratio_rhop_min > 0.0) || !(ratio_rhop_max > ratio_rhop_min) || !(ratio_dtmax > 1.0))
        return fail("or_set_limits: need 0 < ratio_rhop_min < ratio_rhop_max and ratio_dtmax > 1");
    g->lim[0] = ratio_rhop_min;
    g->lim[1] = ratio_rhop_max;
    g->lim[2] = ratio_p_break;
    g->lim[3] = ratio_dtmax;
    if (g->res) set_par(&g->P, g->P.mach, g->P.alpha, g->P.cfl_start, g->lim); /* already preprocessed */
    return 0;
}

int or_counts(or_grid* g, long long* dst, int n, int reset) {
    if (n != OR_NCOUNT) return fail("or_counts: n must be OR_NCOUNT");
    memcpy(dst, g->cnt, sizeof(g->cnt));
    if (reset) memset(g->cnt, 0, sizeof(g->cnt));
    return 0;
}

void or_derived(double mach, double alpha, double d[14]) {
    Par P;
    set_par(&P, mach, alpha, 1.0, NULL);
    const double gam = P.gam, gamm1 = P.gamm1;
    double temp = 1.0 + 0.5 * gamm1 * mach * mach;
    double p_stag = (pow(temp, gam / gamm1) - 1.0) * P.p_bar;
    double e_stag = P.p_bar / gamm1 + 0.5 * P.rho_s * (P.u_s * P.u_s + P.v_s * P.v_s + P.w_s * P.w_s);
    double out[14] = {P.u_s, P.v_s, P.w_s, P.rho_s, P.p_bar, p_stag, e_stag, P.p_min, P.p_max,
                      P.rho_min, P.rho_max, P.p_break, 1.0e5 * e_stag, gamm1};
    memcpy(d, out, sizeof(out));
}

void or_roe_flux(const double ql[5], const double qr[5], double nx, double ny, double nz, double area, double mach,
                 double alpha, double f[5]) {
    Par P;
    set_par(&P, mach, alpha, 1.0, NULL);
    roe(ql, qr, nx, ny, nz, area, &P, f);
}

void or_flux_lusgs(const double q[5], const double dq[5], double nx, double ny, double nz, double mach, double alpha,
                   double f[5]) {
    Par P;
    set_par(&P, mach, alpha, 1.0, NULL);
    flux_lusgs(q, dq, nx, ny, nz, &P, f);
}

int or_sizes(const or_grid* g, int* nC, int* nF, int* nB, int* nN) {
    *nC = g->nC;
    *nF = g->nF;
    *nB = g->nB;
    *nN = g->nN;
    return 0;
}

int or_get(const or_grid* g, const char* name, double* dst, size_t count) {
    const size_t nC = g->nC, n = (size_t)g->nC + g->nB, nF = g->nF;
    const double* src = NULL;
    size_t want = 0;
    const char* st[5] = {"rho", "u", "v", "w", "p"};
    for (int j = 0; j < 5; ++j)
        if (!strcmp(name, st[j])) src = g->q[j], want = n;
    if (!strcmp(name, "res")) src = g->res, want = 5 * nC;
    if (!strcmp(name, "dt")) src = g->dt, want = nC;
    if (!strcmp(name, "dq")) src = g->dq, want = 5 * n;
    if (!strcmp(name, "dqdx")) src = g->dqdx, want = 5 * n;
    if (!strcmp(name, "dqdy")) src = g->dqdy, want = 5 * n;
    if (!strcmp(name, "dqdz")) src = g->dqdz, want = 5 * n;
    if (!strcmp(name, "xfn")) src = g->xfn, want = nF;
    if (!strcmp(name, "yfn")) src = g->yfn, want = nF;
    if (!strcmp(name, "zfn")) src = g->zfn, want = nF;
    if (!strcmp(name, "area")) src = g->area, want = nF;
    if (!strcmp(name, "xfc")) src = g->xfc, want = nF;
    if (!strcmp(name, "yfc")) src = g->yfc, want = nF;
    if (!strcmp(name, "zfc")) src = g->zfc, want = nF;
    if (!strcmp(name, "xcc")) src = g->xcc, want = n;
    if (!strcmp(name, "ycc")) src = g->ycc, want = n;
    if (!strcmp(name, "zcc")) src = g->zcc, want = n;
    if (!strcmp(name, "vol")) src = g->vol, want = n;
    if (!src) return fail("unknown array");
    if (count != want) return fail("wrong count");
    memcpy(dst, src, sizeof(double) * count);
    return 0;
}

void or_free(or_grid* g) {
    if (!g) return;
    void* ptrs[] = {g->x, g->y, g->z, g->nNPF, g->fo, g->f2n, g->f2c, g->btype, g->xfn, g->yfn, g->zfn, g->area,
                    g->xfc, g->yfc, g->zfc, g->xcc, g->ycc, g->zcc, g->vol, g->res, g->dt, g->dq, g->dqdx, g->dqdy,
                    g->dqdz, g->cfo, g->cf, g->cno, g->cn, g->Nmark, g->WeightNode, g->wbf, g->wcn, g->flev_off, g->symm_node, g->symm_n,
                    g->flev_cells};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    for (int j = 0; j < 5; ++j) free(g->q[j]);
    free(g);
}
