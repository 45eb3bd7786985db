// Host set-up of the solver path (PolyGrid after PreprocessGrid).
//
// The metric routines keep the reference's loop structure and expression
// order so the geometry is bitwise equal to the reference's: they are
// per-face accumulations whose per-cell summation order is the face order.
#include "mfs_grid.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace mfs {

void params_default(mfs_params& p) {
    // para_field_global.cpp:37-146
    p = mfs_params{};
    p.mach = 0.8395;
    p.alpha = 3.06;
    p.beita = 0.0;
    p.re = 1.8140449177e7;
    p.T = 288.15;
    p.gam = 1.4;
    p.gascon = 287.053;
    p.cfl_start = 100.0;
    p.cfl_end = 100.0;
    p.cfl_nstep = 2000;
    p.epsa_r = 0.1;
    p.lhs_omga = 1.0;
    p.ratio_dtmax = 1.0e80;
    p.ratio_rhop_max = 10.0;  // :200-202
    p.ratio_rhop_min = 1.0e-4;
    p.ratio_p_break = 0.01;
    params_derive(p);
}

void params_derive(mfs_params& p) {
    // para_field_global.cpp:150-209, same formulas and evaluation order.
    const double tref = 288.15, sref = 110.4, amuref = 1.78938e-5;
    const double trat = (p.T / tref);
    p.amu = amuref * trat * sqrt(trat) * (tref + sref) / (p.T + sref);
    p.ainf = sqrt(p.gam * p.gascon * p.T);
    p.uqq = p.ainf * p.mach;
    const double alpha_t = (p.alpha * kPI / 180.);
    const double beita_t = (p.beita * kPI / 180.);
    p.u_s = p.uqq * cos(alpha_t) * cos(-1. * beita_t);
    p.v_s = p.uqq * sin(-1. * beita_t);
    p.w_s = p.uqq * sin(alpha_t) * cos(-1. * beita_t);
    p.p_s = 0.00;
    p.rho_s = p.re * p.amu / p.uqq;
    p.p_bar = p.rho_s * p.gascon * p.T;
    p.gamm1 = p.gam - 1.0;
    const double temp = 1.0 + 0.5 * p.gamm1 * p.mach * p.mach;
    const double p_stag_t = pow(temp, p.gam / p.gamm1);
    const double rho_stag = pow(temp, 1.0 / p.gamm1) * p.rho_s;
    p.p_stag = (p_stag_t - 1.0) * p.p_bar;
    p.e_stag = p.p_bar / p.gamm1 + 0.5 * p.rho_s * (p.u_s * p.u_s + p.v_s * p.v_s + p.w_s * p.w_s);
    p.p_min = p.ratio_rhop_min * p.p_bar - p.p_bar;
    p.p_max = p.ratio_rhop_max * (p.p_stag + p.p_bar) - p.p_bar;
    p.rho_min = p.ratio_rhop_min * p.rho_s;
    p.rho_max = p.ratio_rhop_max * rho_stag;
    p.p_break = p.ratio_p_break * p.p_bar - p.p_bar;
    p.e_stag_max = 1.0e5 * p.e_stag;
    p.alf_l = p.epsa_r;
    p.alf_n = p.epsa_r;
}

std::string Grid::decouple(int nbParts, int partSize) {
    if (preprocessed) return "Decoupling must run before PreprocessGrid";
    if (decoupled) return "";
    if (nbParts < 2 || nbParts > 4 || partSize < 8)
        return "TDTree parameters out of range (2 <= nbParts <= 4, partSize >= 8)";
    tree = build_tdtree(m, nbParts, partSize);
    m = renumber(m, tree);
    decoupled = true;
    return "";
}

// SpecifyBC (load_grid.cpp:857-902): boundary face k gets the record of its
// patch and ghost cell nTCell + k.
std::string specify_bc(Grid& g) {
    RawMesh& m = g.m;
    g.bctype.assign(m.nB, 0);
    for (int k = 0; k < m.nB; ++k) {
        const int patch = -m.f2c[2 * k + 1];
        const int rec = patch - 1;
        if (rec < 0 || rec >= (int)m.bc_type.size()) {
            std::ostringstream os;
            os << "boundary face " << k << " names patch " << patch << " but bc.para has " << m.bc_type.size() << " records";
            return os.str();
        }
        g.bctype[k] = m.bc_type[rec];
        m.f2c[2 * k + 1] = k + m.nC;
    }
    return "";
}

// ComputeMetrics (grid_polyhedra.cpp:1071-1100) and its five steps
// (grid_polyhedra_auxiliary_functions.cpp:68-808).
double centroid_scale() {
    const int iexp = 16;  // CorrectCellCentroid (:679-690)
    return pow(10.0, -iexp + 1);
}

std::string compute_metrics(Grid& g) {
    const RawMesh& m = g.m;
    const int nC = m.nC, nB = m.nB, nF = m.nF, n = nC + nB;
    const int* f2c = m.f2c.data();
    const int* f2n = m.f2n.data();
    const int* nNPF = m.nNPF.data();
    const double *x = m.x.data(), *y = m.y.data(), *z = m.z.data();
    g.vol.assign(n, 0.0);
    g.xcc.assign(n, 0.0);
    g.ycc.assign(n, 0.0);
    g.zcc.assign(n, 0.0);
    g.area.assign(nF, 0.0);
    g.xfn.assign(nF, 0.0);
    g.yfn.assign(nF, 0.0);
    g.zfn.assign(nF, 0.0);
    g.xfc.assign(nF, 0.0);
    g.yfc.assign(nF, 0.0);
    g.zfc.assign(nF, 0.0);
    double *xfc = g.xfc.data(), *yfc = g.yfc.data(), *zfc = g.zfc.data();
    double *xcc = g.xcc.data(), *ycc = g.ycc.data(), *zcc = g.zcc.data();
    double *xfn = g.xfn.data(), *yfn = g.yfn.data(), *zfn = g.zfn.data();
    double *area = g.area.data(), *vol = g.vol.data();

    // FaceCellCenterbyAverage (:271-355)
    {
        int node = 0;
        for (int i = 0; i < nF; i++) {
            xfc[i] = 0.0;
            yfc[i] = 0.0;
            zfc[i] = 0.0;
            for (int j = 0; j < nNPF[i]; j++) {
                const int p1 = f2n[node++];
                xfc[i] += x[p1];
                yfc[i] += y[p1];
                zfc[i] += z[p1];
            }
            xfc[i] /= nNPF[i];
            yfc[i] /= nNPF[i];
            zfc[i] /= nNPF[i];
        }
        std::vector<int> nface(nC, 0);
        int cell = 0;
        for (int i = 0; i < nB; i++) {
            const int c1 = f2c[cell++];
            cell++;
            xcc[c1] += xfc[i];
            ycc[c1] += yfc[i];
            zcc[c1] += zfc[i];
            nface[c1]++;
        }
        for (int i = nB; i < nF; i++) {
            const int c1 = f2c[cell++];
            const int c2 = f2c[cell++];
            xcc[c1] += xfc[i];
            ycc[c1] += yfc[i];
            zcc[c1] += zfc[i];
            xcc[c2] += xfc[i];
            ycc[c2] += yfc[i];
            zcc[c2] += zfc[i];
            nface[c1]++;
            nface[c2]++;
        }
        for (int i = 0; i < nC; i++) {
            xcc[i] /= nface[i];
            ycc[i] /= nface[i];
            zcc[i] /= nface[i];
        }
    }

    // FaceAreaNormalCentroid_cycle (:68-266): 4 fixed-point passes on the
    // face centroid, then normals and areas.
    {
        std::vector<double> xfct(nF), yfct(nF), zfct(nF);
        for (int k = 0; k < 4; k++) {
            int count = 0;
            for (int i = 0; i < nF; i++) {
                area[i] = 0.0;
                xfct[i] = 0.0;
                yfct[i] = 0.0;
                zfct[i] = 0.0;
                const int p1 = f2n[count];
                for (int j = 0; j < nNPF[i]; j++) {
                    const int p2 = f2n[count++];
                    const int p3 = (j == nNPF[i] - 1) ? p1 : f2n[count];
                    const double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
                    const double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
                    const double xn = y21 * z31 - y31 * z21;
                    const double yn = z21 * x31 - z31 * x21;
                    const double zn = x21 * y31 - x31 * y21;
                    const double art = sqrt(xn * xn + yn * yn + zn * zn);
                    const double xt = x[p2] + x[p3] + xfc[i];
                    const double yt = y[p2] + y[p3] + yfc[i];
                    const double zt = z[p2] + z[p3] + zfc[i];
                    area[i] += art;
                    xfct[i] += art * xt;
                    yfct[i] += art * yt;
                    zfct[i] += art * zt;
                }
                if (area[i] > kTINY) {
                    const double art = 1.0 / area[i] / 3.0;
                    xfct[i] *= art;
                    yfct[i] *= art;
                    zfct[i] *= art;
                    xfc[i] = xfct[i];
                    yfc[i] = yfct[i];
                    zfc[i] = zfct[i];
                }
            }
        }
        int count = 0;
        for (int i = 0; i < nF; i++) {
            xfn[i] = 0.0;
            yfn[i] = 0.0;
            zfn[i] = 0.0;
            area[i] = 0.0;
            xfct[i] = 0.0;
            yfct[i] = 0.0;
            zfct[i] = 0.0;
            const int p1 = f2n[count];
            for (int j = 0; j < nNPF[i]; j++) {
                const int p2 = f2n[count++];
                const int p3 = (j == nNPF[i] - 1) ? p1 : f2n[count];
                const double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
                const double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
                const double xn = y21 * z31 - y31 * z21;
                const double yn = z21 * x31 - z31 * x21;
                const double zn = x21 * y31 - x31 * y21;
                const double art = sqrt(xn * xn + yn * yn + zn * zn);
                const double xt = x[p2] + x[p3] + xfc[i];
                const double yt = y[p2] + y[p3] + yfc[i];
                const double zt = z[p2] + z[p3] + zfc[i];
                area[i] += art;
                xfct[i] += art * xt;
                yfct[i] += art * yt;
                zfct[i] += art * zt;
                xfn[i] += xn;
                yfn[i] += yn;
                zfn[i] += zn;
            }
            if (area[i] > kTINY) {
                const double art = 1.0 / area[i] / 3.0;
                xfc[i] = xfct[i] * art;
                yfc[i] = yfct[i] * art;
                zfc[i] = zfct[i] * art;
            }
            const double art = sqrt(xfn[i] * xfn[i] + yfn[i] * yfn[i] + zfn[i] * zfn[i]);
            area[i] = art * 0.5;
            if (art > kTINY) {
                xfn[i] /= art;
                yfn[i] /= art;
                zfn[i] /= art;
            } else {
                const int c1 = f2c[i + i];
                const double xn = xfc[i] - xcc[c1], yn = yfc[i] - ycc[c1], zn = zfc[i] - zcc[c1];
                const double a2 = sqrt(xn * xn + yn * yn + zn * zn);
                xfn[i] = xn / a2;
                yfn[i] = yn / a2;
                zfn[i] = zn / a2;
            }
        }
    }

    // CellVolCentroid (:357-612)
    {
        std::vector<double> xcct(nC, 0.0), ycct(nC, 0.0), zcct(nC, 0.0);
        for (int i = 0; i < nC; i++) vol[i] = 0.0;
        int cell = 0, node = 0;
        for (int i = 0; i < nB; i++) {
            const int c1 = f2c[cell++];
            cell++;
            const int p1 = f2n[node];
            for (int j = 0; j < nNPF[i]; j++) {
                const int p2 = f2n[node++];
                const int p3 = (j == nNPF[i] - 1) ? p1 : f2n[node];
                const double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
                const double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
                const double nx = y21 * z31 - y31 * z21;
                const double ny = z21 * x31 - z31 * x21;
                const double nz = x21 * y31 - x31 * y21;
                const double x41 = xcc[c1] - xfc[i], y41 = ycc[c1] - yfc[i], z41 = zcc[c1] - zfc[i];
                const double xt = x[p2] + x[p3] + xfc[i] + xcc[c1];
                const double yt = y[p2] + y[p3] + yfc[i] + ycc[c1];
                const double zt = z[p2] + z[p3] + zfc[i] + zcc[c1];
                double volt = nx * x41 + ny * y41 + nz * z41;
                volt = -volt;
                xcct[c1] += volt * xt;
                ycct[c1] += volt * yt;
                zcct[c1] += volt * zt;
                vol[c1] += volt;
            }
        }
        for (int i = nB; i < nF; i++) {
            const int c1 = f2c[cell++];
            const int c2 = f2c[cell++];
            const int p1 = f2n[node];
            for (int j = 0; j < nNPF[i]; j++) {
                const int p2 = f2n[node++];
                const int p3 = (j == nNPF[i] - 1) ? p1 : f2n[node];
                const double x21 = x[p2] - xfc[i], y21 = y[p2] - yfc[i], z21 = z[p2] - zfc[i];
                const double x31 = x[p3] - xfc[i], y31 = y[p3] - yfc[i], z31 = z[p3] - zfc[i];
                const double nx = y21 * z31 - y31 * z21;
                const double ny = z21 * x31 - z31 * x21;
                const double nz = x21 * y31 - x31 * y21;
                double x41 = xcc[c1] - xfc[i], y41 = ycc[c1] - yfc[i], z41 = zcc[c1] - zfc[i];
                double xt = x[p2] + x[p3] + xfc[i] + xcc[c1];
                double yt = y[p2] + y[p3] + yfc[i] + ycc[c1];
                double zt = z[p2] + z[p3] + zfc[i] + zcc[c1];
                double volt = nx * x41 + ny * y41 + nz * z41;
                volt = -volt;
                xcct[c1] += volt * xt;
                ycct[c1] += volt * yt;
                zcct[c1] += volt * zt;
                vol[c1] += volt;
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
        int small = 0;
        for (int i = 0; i < nC; i++) {
            const double tmp = 1.0 / (4.0 * vol[i] + kTINY);
            xcct[i] *= tmp;
            ycct[i] *= tmp;
            zcct[i] *= tmp;
            vol[i] /= 6.0;
            if (vol[i] < kTINY) ++small;
        }
        if (small > 0) {
            std::ostringstream os;
            os << "There are " << small << " cell's vol too small! This must correct!";
            return os.str();  // the reference prints this and calls exit(1) (:528-536)
        }
        std::vector<char> mark(nC, 1);
        for (int i = 0; i < nB; i++) {
            const int c1 = f2c[i + i];
            const double x21 = xcct[c1] - xfc[i], y21 = ycct[c1] - yfc[i], z21 = zcct[c1] - zfc[i];
            const double dot = x21 * xfn[i] + y21 * yfn[i] + z21 * zfn[i];
            if (dot > 0.0) mark[c1] = 0;
        }
        for (int i = nB; i < nF; i++) {
            const int c1 = f2c[i + i], c2 = f2c[i + i + 1];
            double x21 = xcct[c1] - xfc[i], y21 = ycct[c1] - yfc[i], z21 = zcct[c1] - zfc[i];
            double dot = x21 * xfn[i] + y21 * yfn[i] + z21 * zfn[i];
            if (dot > 0.0) mark[c1] = 0;
            x21 = xcct[c2] - xfc[i];
            y21 = ycct[c2] - yfc[i];
            z21 = zcct[c2] - zfc[i];
            dot = x21 * xfn[i] + y21 * yfn[i] + z21 * zfn[i];
            if (dot < 0.0) mark[c2] = 0;
        }
        for (int i = 0; i < nC; i++)
            if (mark[i] == 1) {
                xcc[i] = xcct[i];
                ycc[i] = ycct[i];
                zcc[i] = zcct[i];
            }
        cell = 0;
        for (int i = 0; i < nB; i++) {
            const int c1 = f2c[cell++];
            cell++;
            const int c2 = i + nC;
            if (area[i] > kTINY) {
                const double tmp =
                    2. * ((xcc[c1] - xfc[i]) * xfn[i] + (ycc[c1] - yfc[i]) * yfn[i] + (zcc[c1] - zfc[i]) * zfn[i]);
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
    }

    // CorrectFaceNormal (:619-671)
    {
        int node = 0;
        for (int i = 0; i < nF; i++) {
            const int p1 = f2n[node];
            node += nNPF[i];
            if (area[i] < kTINY) {
                const int c1 = f2c[i + i];
                double x21 = xcc[c1] - xfc[i], y21 = ycc[c1] - yfc[i], z21 = zcc[c1] - zfc[i];
                const double x31 = x[p1] - xfc[i], y31 = y[p1] - yfc[i], z31 = z[p1] - zfc[i];
                const double dot = x21 * x31 + y21 * y31 + z21 * z31;
                double nn = x31 * x31 + y31 * y31 + z31 * z31;
                double ratio = 0.0;
                if (nn > kTINY) ratio = dot / nn;
                x21 -= ratio * x31;
                y21 -= ratio * y31;
                z21 -= ratio * z31;
                nn = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
                xfn[i] = -x21 / nn;
                yfn[i] = -y21 / nn;
                zfn[i] = -z21 / nn;
            }
        }
    }

    // CorrectCellCentroid (:679-808)
    {
        const double scaleXYZcc = centroid_scale();
        const double minDn = 1.0e-10;
        auto absmax_signfirst = [](double a, double b) {  // AbsMaxSignFirst (:812-823)
            const int sign = (a > 0) ? 1 : -1;
            double c = std::max(fabs(a), fabs(b));
            c *= sign;
            return c;
        };
        auto correct = [&](int c, int i) {
            const double x21 = xcc[c] - xfc[i], y21 = ycc[c] - yfc[i], z21 = zcc[c] - zfc[i];
            const double dot = x21 * xfn[i] + y21 * yfn[i] + z21 * zfn[i];
            if (fabs(dot) < minDn) {
                double dx = minDn * xfn[i], dy = minDn * yfn[i], dz = minDn * zfn[i];
                dx = absmax_signfirst(dx, scaleXYZcc * xcc[c]);
                dy = absmax_signfirst(dy, scaleXYZcc * ycc[c]);
                dz = absmax_signfirst(dz, scaleXYZcc * zcc[c]);
                const int sign = (dot > 0.0) ? 1 : -1;
                xcc[c] += sign * dx;
                ycc[c] += sign * dy;
                zcc[c] += sign * dz;
            }
        };
        for (int i = 0; i < nB; ++i) correct(f2c[i + i], i);
        for (int i = nB; i < nF; i++) {
            correct(f2c[i + i], i);
            correct(f2c[i + i + 1], i);
        }
    }
    return "";
}

// AllocAndInitFlowfield (load_grid.cpp:571-601): uniform freestream, gauge p.
void init_flowfield(Grid& g) {
    const int n = g.m.nC + g.m.nB;
    g.rho.assign(n, g.params.rho_s);
    g.u.assign(n, g.params.u_s);
    g.v.assign(n, g.params.v_s);
    g.w.assign(n, g.params.w_s);
    g.p.assign(n, g.params.p_s);
}

// SetGhostVariables (load_grid.cpp:953-1212), inviscid branch.
void set_ghost_variables(Grid& g) {
    const mfs_params& P = g.params;
    const int nB = g.m.nB;
    const int* f2c = g.m.f2c.data();
    double *rho = g.rho.data(), *u = g.u.data(), *v = g.v.data(), *w = g.w.data(), *p = g.p.data();
    const double *xfn = g.xfn.data(), *yfn = g.yfn.data(), *zfn = g.zfn.data();
    const double gam = P.gam, p_bar = P.p_bar;
    const double rho00 = P.rho_s, u00 = P.u_s, v00 = P.v_s, w00 = P.w_s, p00 = P.p_s;
    const double norm_of_uvw = sqrt(u00 * u00 + v00 * v00 + w00 * w00);
    const double eps_of_farfield_vn = 0.0;
    int count = 0;
    for (int i = 0; i < nB; i++) {
        const int type = g.bctype[i];
        const int c1 = f2c[count++];
        const int c2 = f2c[count++];
        if (type == BC_INTERFACE) continue;
        double vn;
        switch (type) {
            case BC_WALL:
                p[c2] = p[c1];
                rho[c2] = rho[c1];
                vn = 2. * (xfn[i] * u[c1] + yfn[i] * v[c1] + zfn[i] * w[c1]);
                u[c2] = u[c1] - vn * xfn[i];
                v[c2] = v[c1] - vn * yfn[i];
                w[c2] = w[c1] - vn * zfn[i];
                break;
            case BC_SYMM:
                rho[c2] = rho[c1];
                p[c2] = p[c1];
                vn = 2. * (xfn[i] * u[c1] + yfn[i] * v[c1] + zfn[i] * w[c1]);
                u[c2] = u[c1] - vn * xfn[i];
                v[c2] = v[c1] - vn * yfn[i];
                w[c2] = w[c1] - vn * zfn[i];
                break;
            case BC_FAR_FIELD: {
                const double um = u00, vm = v00, wm = w00;
                const double up = u[c1], vp = v[c1], wp = w[c1];
                const double rhom = rho00, pm = p00 + p_bar;
                const double rhop = rho[c1], pp = p[c1] + p_bar;
                const double vnm = xfn[i] * um + yfn[i] * vm + zfn[i] * wm;
                const double vnp = xfn[i] * up + yfn[i] * vp + zfn[i] * wp;
                const double cm = sqrt(gam * pm / rhom);
                const double cp = sqrt(gam * pp / rhop);
                const double riemm = vnm - 2. * cm / (gam - 1.);
                const double riemp = vnp + 2. * cp / (gam - 1.);
                const double vnb = 0.5 * (riemp + riemm);
                const double cb = 0.25 * (riemp - riemm) * (gam - 1.);
                if (fabs(vnb / cb) > 1.) {
                    if (vnb <= 0.0) {
                        rho[c2] = rhom;
                        u[c2] = um;
                        v[c2] = vm;
                        w[c2] = wm;
                        p[c2] = pm;
                    } else {
                        rho[c2] = rhop;
                        u[c2] = up;
                        v[c2] = vp;
                        w[c2] = wp;
                        p[c2] = pp;
                    }
                } else {
                    const double rela_vnb = vnb / norm_of_uvw;
                    if (rela_vnb <= -eps_of_farfield_vn) {
                        const double vtx = um - vnm * xfn[i], vty = vm - vnm * yfn[i], vtz = wm - vnm * zfn[i];
                        const double entr = pm / pow(rhom, gam);
                        rho[c2] = pow((cb * cb / (entr * gam)), double(1. / (gam - 1.)));
                        u[c2] = vtx + vnb * xfn[i];
                        v[c2] = vty + vnb * yfn[i];
                        w[c2] = vtz + vnb * zfn[i];
                        p[c2] = cb * cb * rho[c2] / gam;
                    } else if (rela_vnb > eps_of_farfield_vn) {
                        const double vtx = up - vnp * xfn[i], vty = vp - vnp * yfn[i], vtz = wp - vnp * zfn[i];
                        const double entr = pp / pow(rhop, gam);
                        rho[c2] = pow((cb * cb / (entr * gam)), double(1. / (gam - 1.)));
                        u[c2] = vtx + vnb * xfn[i];
                        v[c2] = vty + vnb * yfn[i];
                        w[c2] = vtz + vnb * zfn[i];
                        p[c2] = cb * cb * rho[c2] / gam;
                    } else {
                        rho[c2] = 0.5 * (rhop + rhom);
                        u[c2] = 0.5 * (up + um);
                        v[c2] = 0.5 * (vp + vm);
                        w[c2] = 0.5 * (wp + wm);
                        p[c2] = 0.5 * (pp + pm);
                    }
                }
                rho[c2] = 2 * rho[c2] - rhop;
                u[c2] = 2 * u[c2] - up;
                v[c2] = 2 * v[c2] - vp;
                w[c2] = 2 * w[c2] - wp;
                p[c2] = 2 * p[c2] - pp;
                p[c2] -= p_bar;
                break;
            }
            default:
                break;  // the reference prints "Error in SetGhostVariables 001" and continues
        }
        rho[c2] = std::max(rho[c2], P.rho_min);
        rho[c2] = std::min(rho[c2], P.rho_max);
        p[c2] = std::max(p[c2], P.p_min);
        p[c2] = std::min(p[c2], P.p_max);
    }
}

std::string Grid::preprocess(const mfs_params& prm, bool device_metrics) {
    if (preprocessed) return "grid already preprocessed";
    params = prm;
    // SpecifyBC renames the boundary faces' second cells (patch -> ghost);
    // a failure after it puts them back, so that the grid can be retried
    const std::vector<int> f2c_boundary(m.f2c.begin(), m.f2c.begin() + 2 * (size_t)m.nB);
    std::string e = specify_bc(*this);
    if (e.empty()) e = device_metrics ? compute_metrics_device(*this) : compute_metrics(*this);
    if (!e.empty()) {
        std::copy(f2c_boundary.begin(), f2c_boundary.end(), m.f2c.begin());
        return e;
    }
    init_flowfield(*this);
    set_ghost_variables(*this);
    preprocessed = true;
    return "";
}

}  // namespace mfs
