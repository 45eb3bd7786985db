// Point-wise physics of the solver path, shared by the HIP kernels.
//
// Each function keeps the reference's expression trees operand for operand
// (evaluation order, association, divisions kept as divisions) so that,
// compiled without floating-point contraction, it rounds exactly like the
// reference's serial code.
#pragma once
#include <hip/hip_runtime.h>

namespace mfs {

#define MFS_HD __host__ __device__ __forceinline__

constexpr double kBIGd = 1.e40;   // BIG (src/Base/number_type.h:28)
constexpr double kTINYd = 1.e-40; // TINY (src/Base/number_type.h:29)

// std::max / std::min semantics (the reference's MAX/MIN macros,
// src/Base/constant.h:55-60).
template <typename T>
MFS_HD T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
MFS_HD T smin(T a, T b) { return (b < a) ? b : a; }

struct FlowConst {
    double gam, gamm1, p_bar, alf_l, lhs_omga;
    double rho_min, rho_max, p_min, p_max, p_break;
    double alpq, phiq, betq;  // CFL3D under-relaxation (forward_lusgs.cpp:1376-1378)
};

// RoeFlux_noprec live loop body (src/Kernels/cal_flux.cpp:682-833):
// flux of face (ql -> qr) with unit normal (ax,ay,az) and area.
template <typename R>
MFS_HD void roe_flux(const R ql[5], const R qr[5], R ax, R ay, R az, R area, R gamm1, R p_bar, R alf_l, R tiny,
                     R f[5]) {
    R et_l = (ql[4] + p_bar) / gamm1 + R(0.5) * ql[0] * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
    R et_r = (qr[4] + p_bar) / gamm1 + R(0.5) * qr[0] * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    R ht_l = et_l + ql[4] + p_bar;
    R ht_r = et_r + qr[4] + p_bar;
    R vn_l = ax * ql[1] + ay * ql[2] + az * ql[3];
    R vn_r = ax * qr[1] + ay * qr[2] + az * qr[3];
    R tmp0 = vn_l * ql[0];
    R tmp1 = vn_r * qr[0];
    f[0] = tmp0 + tmp1;
    f[1] = tmp0 * ql[1] + ax * ql[4] + tmp1 * qr[1] + ax * qr[4];
    f[2] = tmp0 * ql[2] + ay * ql[4] + tmp1 * qr[2] + ay * qr[4];
    f[3] = tmp0 * ql[3] + az * ql[4] + tmp1 * qr[3] + az * qr[4];
    f[4] = ht_l * vn_l + ht_r * vn_r;

    tmp0 = sqrt(qr[0] / ql[0]);
    tmp1 = R(1.0) / (R(1.0) + tmp0);
    R rho_a = sqrt(qr[0] * ql[0]);
    R u_a = (ql[1] + qr[1] * tmp0) * tmp1;
    R v_a = (ql[2] + qr[2] * tmp0) * tmp1;
    R w_a = (ql[3] + qr[3] * tmp0) * tmp1;
    R vn_a = u_a * ax + v_a * ay + w_a * az;
    R h_a = (ht_l / ql[0] + ht_r / qr[0] * tmp0) * tmp1;
    R q2 = R(0.5) * (u_a * u_a + v_a * v_a + w_a * w_a);
    R c2_a = gamm1 * (h_a - q2);
    c2_a = fabs(c2_a);
    R c_a = sqrt(c2_a);
    R eigv1 = fabs(vn_a);
    R eigv2 = fabs(vn_a + c_a);
    R eigv3 = fabs(vn_a - c_a);

    R epsa_r = alf_l;
    R spectral = fabs(u_a) + fabs(v_a) + fabs(w_a) + c_a;
    R epsaa = epsa_r * spectral;
    R epsbb = R(0.25) / smax(epsaa, tiny);
    R epscc = R(2.0) * epsaa;
    if (eigv1 < epscc) eigv1 = eigv1 * eigv1 * epsbb + epsaa;
    if (eigv2 < epscc) eigv2 = eigv2 * eigv2 * epsbb + epsaa;
    if (eigv3 < epscc) eigv3 = eigv3 * eigv3 * epsbb + epsaa;

    R drho = qr[0] - ql[0];
    R du = qr[1] - ql[1];
    R dv = qr[2] - ql[2];
    R dw = qr[3] - ql[3];
    R dp = qr[4] - ql[4];
    R dvn = vn_r - vn_l;
    R dq2 = u_a * du + v_a * dv + w_a * dw;

    tmp0 = dp / c2_a;
    tmp1 = rho_a * dvn / c_a;
    R alpha1 = (drho - tmp0) * eigv1;
    R alpha2 = R(0.5) * (tmp0 + tmp1) * eigv2;
    R alpha3 = R(0.5) * (tmp0 - tmp1) * eigv3;

    tmp0 = alpha1 + alpha2 + alpha3;
    tmp1 = eigv1 * rho_a;
    R tmp2 = -tmp1 * dvn + (alpha2 - alpha3) * c_a;
    f[0] -= tmp0;
    f[1] -= tmp0 * u_a + tmp1 * du + tmp2 * ax;
    f[2] -= tmp0 * v_a + tmp1 * dv + tmp2 * ay;
    f[3] -= tmp0 * w_a + tmp1 * dw + tmp2 * az;
    f[4] -= alpha1 * q2 + (alpha2 + alpha3) * h_a + tmp1 * dq2 + tmp2 * vn_a;

    tmp0 = R(0.5) * area;
    f[0] *= tmp0;
    f[1] *= tmp0;
    f[2] *= tmp0;
    f[3] *= tmp0;
    f[4] *= tmp0;
}

// FluxLUSGS3D (src/Kernels/forward_lusgs.cpp:1271-1324): neighbour state q,
// neighbour increment DQ, normal n pointing from the swept cell to the
// neighbour.
// The part of FluxLUSGS3D that does not depend on DQ (the neighbour's
// state, its flux and the spectral radius), so that a sweep can evaluate
// it before the neighbour's DQ arrives.  flux_lusgs_pre followed by
// flux_lusgs_post performs the reference's operations in its order.
template <typename R>
struct FluxPre {
    R Q[5];          // conservative state of the neighbour
    R rho, v_n, eig;  // density, normal velocity, lhs_omga * spectral radius
    R peff, gam1;
    R f[5];          // -F(Q) (components 1-4)
};

template <typename R>
MFS_HD void flux_lusgs_pre(FluxPre<R>& P, const R q[5], R nx, R ny, R nz, R gam, R p_bar, R lhs_omga) {
    R rv2, p, c2;
    P.gam1 = gam - R(1.);
    P.Q[0] = q[0];
    P.Q[1] = q[0] * q[1];
    P.Q[2] = q[0] * q[2];
    P.Q[3] = q[0] * q[3];
    rv2 = R(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    p = q[4];
    P.Q[4] = p / P.gam1 + rv2;

    P.v_n = q[1] * nx + q[2] * ny + q[3] * nz;
    c2 = gam * (p + p_bar) / q[0];
    P.eig = fabs(P.v_n) + sqrt(c2);
    P.eig *= lhs_omga;

    P.peff = gam * p_bar / P.gam1;
    P.f[1] = -P.Q[1] * P.v_n - p * nx;
    P.f[2] = -P.Q[2] * P.v_n - p * ny;
    P.f[3] = -P.Q[3] * P.v_n - p * nz;
    P.f[4] = -(P.Q[4] + p + P.peff) * P.v_n;
    P.rho = q[0];
}

template <typename R>
MFS_HD void flux_lusgs_post(const FluxPre<R>& P, const R DQ[5], R nx, R ny, R nz, R flux[5]) {
    R Q[5], rv2, p, v_n;
    for (int i = 0; i < 5; i++) Q[i] = P.Q[i] + DQ[i];
    rv2 = R(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) / Q[0];
    p = P.gam1 * (Q[4] - rv2);

    flux[0] = DQ[1] * nx + DQ[2] * ny + DQ[3] * nz;
    v_n = P.v_n * P.rho;
    v_n += flux[0];
    v_n /= Q[0];
    flux[1] = P.f[1] + (Q[1] * v_n + p * nx);
    flux[2] = P.f[2] + (Q[2] * v_n + p * ny);
    flux[3] = P.f[3] + (Q[3] * v_n + p * nz);
    flux[4] = P.f[4] + (Q[4] + p + P.peff) * v_n;

    for (int i = 0; i < 5; i++) flux[i] -= P.eig * DQ[i];
}

template <typename R>
MFS_HD void flux_lusgs(const R q[5], const R DQ[5], R nx, R ny, R nz, R gam, R p_bar, R lhs_omga, R flux[5]) {
    FluxPre<R> P;
    flux_lusgs_pre(P, q, nx, ny, nz, gam, p_bar, lhs_omga);
    flux_lusgs_post(P, DQ, nx, ny, nz, flux);
}

// FluxLUSGS3D in increment form, for the FP32 operator of mixed precision.
// The reference's expression forms F(Q+DQ) and F(Q) separately and
// subtracts them; in FP32 the energy terms ((Q4+p+peff)*v_n, peff =
// gam*p_bar/gam1 ~ 3e5) leave a rounding error of ~1e-7*|F| that does not
// shrink with DQ.  Expanding the difference algebraically gives the same
// function with every term proportional to DQ:
//   dvn = (DQ1..3.n - v_n*DQ0)/Q0',  Q' = Q + DQ
//   drv2 = 0.5*(2*Q1..3.DQ1..3 + |DQ1..3|^2 - Q0*|u|^2*DQ0)/Q0'
//   dp = gam1*(DQ4 - drv2)
//   dF0 = DQ1..3.n,  dF_i = Q_i*dvn + DQ_i*v_n' + dp*n_i,  dF4 = (Q4+p+peff)*dvn + (DQ4+dp)*v_n'
//   flux = dF - eig*DQ
template <typename R>
struct FluxDeltaPre {
    R rho, Q1, Q2, Q3, vv, v_n, eig, gam1;
    R H;  // the energy flux factor Q4 + p + peff
};

template <typename R>
MFS_HD void flux_lusgs_delta_pre(FluxDeltaPre<R>& P, const R q[5], R nx, R ny, R nz, R gam, R p_bar, R lhs_omga) {
    P.gam1 = gam - R(1.);
    const R rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    P.rho = rho;
    P.Q1 = rho * u;
    P.Q2 = rho * v;
    P.Q3 = rho * w;
    P.vv = u * u + v * v + w * w;
    const R Q4 = p / P.gam1 + R(0.5) * rho * P.vv;
    P.v_n = u * nx + v * ny + w * nz;
    P.eig = (fabs(P.v_n) + sqrt(gam * (p + p_bar) / rho)) * lhs_omga;
    const R peff = gam * p_bar / P.gam1;
    P.H = Q4 + p + peff;
}

template <typename R>
MFS_HD void flux_lusgs_delta_post(const FluxDeltaPre<R>& P, const R DQ[5], R nx, R ny, R nz, R flux[5]) {
    const R Q0n = P.rho + DQ[0];
    const R dmn = DQ[1] * nx + DQ[2] * ny + DQ[3] * nz;  // increment of the mass flux
    const R dvn = (dmn - P.v_n * DQ[0]) / Q0n;
    const R vn1 = P.v_n + dvn;
    const R qd = P.Q1 * DQ[1] + P.Q2 * DQ[2] + P.Q3 * DQ[3];
    const R dd = DQ[1] * DQ[1] + DQ[2] * DQ[2] + DQ[3] * DQ[3];
    const R drv2 = R(0.5) * (R(2.) * qd + dd - P.rho * P.vv * DQ[0]) / Q0n;
    const R dp = P.gam1 * (DQ[4] - drv2);
    flux[0] = dmn;
    flux[1] = P.Q1 * dvn + DQ[1] * vn1 + dp * nx;
    flux[2] = P.Q2 * dvn + DQ[2] * vn1 + dp * ny;
    flux[3] = P.Q3 * dvn + DQ[3] * vn1 + dp * nz;
    flux[4] = P.H * dvn + (DQ[4] + dp) * vn1;
    for (int i = 0; i < 5; i++) flux[i] -= P.eig * DQ[i];
}

template <typename R>
MFS_HD void flux_lusgs_delta(const R q[5], const R DQ[5], R nx, R ny, R nz, R gam, R p_bar, R lhs_omga, R flux[5]) {
    FluxDeltaPre<R> P;
    flux_lusgs_delta_pre(P, q, nx, ny, nz, gam, p_bar, lhs_omga);
    flux_lusgs_delta_post(P, DQ, nx, ny, nz, flux);
}

// The three-stage DQ positivity limiter (DQ_limit == 2) of both sweeps
// (forward_lusgs.cpp:939-973 and 1184-1218).  q is the swept cell's state.
template <typename R>
MFS_HD void dq_limit(const R q[5], R DQ[5], R gamm1, R rho_min, R rho_max, R p_min, R p_max) {
    R vv = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    for (int stage = 0; stage < 3; ++stage) {
        R dp = DQ[4] + R(0.5) * DQ[0] * vv - (DQ[1] * q[1] + DQ[2] * q[2] + DQ[3] * q[3]);
        dp *= gamm1;
        if ((q[0] + DQ[0]) < rho_min || (q[0] + DQ[0]) > rho_max || (q[4] + dp) < p_min || (q[4] + dp) > p_max) {
            if (stage < 2) {
                for (int i = 0; i < 5; ++i) DQ[i] *= R(0.1);
            } else {
                for (int i = 0; i < 5; ++i) DQ[i] = R(0.0);
            }
        }
    }
}

// UpdateFlowField3D_CFL3d cell body (forward_lusgs.cpp:1386-1428); q holds
// rho,u,v,w,p (gauge) and is updated in place.
MFS_HD void update_cell(double q[5], const double DQ[5], const FlowConst& K) {
    const double gam1 = K.gam - 1.;
    double rhot = q[0];
    double ru = rhot * q[1];
    double rv = rhot * q[2];
    double rw = rhot * q[3];
    double re = q[4] / gam1 + 0.5 * rhot * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    rhot += DQ[0];
    ru += DQ[1];
    rv += DQ[2];
    rw += DQ[3];
    re += DQ[4];
    double rhotr = 1. / (rhot + kTINYd);
    double un = ru * rhotr, vn = rv * rhotr, wn = rw * rhotr;
    double rho_del = DQ[0];
    double rho_rat = rho_del / q[0];
    if (rho_rat < K.alpq) rho_del /= K.betq + fabs(rho_rat) * K.phiq;
    double rho = q[0] + rho_del;
    rho = smax(rho, K.rho_min);
    rho = smin(rho, K.rho_max);
    double ptmp = gam1 * (re - 0.5 * (un * un + vn * vn + wn * wn) * rho);
    double p_del = ptmp - q[4];
    double p_rat = p_del / (q[4] + K.p_bar);
    if (p_rat < K.alpq) p_del /= K.betq + fabs(p_rat) * K.phiq;
    double p = q[4] + p_del;
    p = smax(p, K.p_min);
    p = smin(p, K.p_max);
    q[0] = rho;
    q[1] = un;
    q[2] = vn;
    q[3] = wn;
    q[4] = p;
}

}  // namespace mfs
