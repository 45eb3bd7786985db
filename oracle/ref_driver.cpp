// Test infrastructure (oracle): a driver main for the reference's own serial
// sources, compiled from /root/reference by oracle/Makefile into oracle/_ref/.
// It replaces src/main.cpp only so the mesh path, flow condition and
// iteration count can be given on the command line, and so the state can be
// dumped after chosen iterations.  The iteration body is the loop of
// src/main.cpp:198-291 (ZeroResiduals, InviscidFlux, ComputeTimeStep,
// ForwardLUSGS, PassInterfaceData, CalculateGradient, iter_done++), and the
// residual norm is the one of src/main.cpp:294-317, evaluated after every
// iteration instead of once at the end.
//
// Nothing in the product loads or runs this program.
//
// Built three ways by oracle/Makefile: the serial variant (ref_driver), the
// OpenMP face-colouring variant of command:22 (ref_driver_omp: -fopenmp
// -DMF_OPENMP -DOMP_CellColor -DOMP_FaceColor, a timing baseline; its
// coloured LU-SGS sums in another order), and with -DMFS_SHIM (shim_driver:
// the reference's set-up and this loop, with the solver functions of
// minifs_amd/shim/mfs_gpu_shim.cpp on libmfs_hip.so in place of the four
// kernel files).
//
// Usage: ref_driver MESH BCFILE OUTDIR MACH ALPHA CFL ITERS DUMPLIST [options]
//   DUMPLIST: comma-separated iteration counts after which the state is
//   written (0 = after setup), e.g. "0,1,10,100".
//   options: ghosts | nogeo | ramp=CFL_START,NSTEP |
//            limits=RHOP_MIN,RHOP_MAX,P_BREAK (ratio_rhop_min, ratio_rhop_max,
//            ratio_p_break of para_field_global.cpp:200-202, applied before
//            the derived limits are computed) | dtratio=X (ratio_dtmax,
//            :115) | init=DIR (after PreprocessGrid, rho u v w p of all
//            nTCell+nBFace cells, ghosts included, are read from
//            DIR/init_<name>.f64)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>

#include "boundary_condition.h"
#include "cal_flux.h"
#include "cal_gradient.h"
#include "compute_timestep.h"
#include "forward_lusgs.h"
#include "grid_polyhedra.h"
#include "load_grid.h"
#include "para_field_global.h"
#ifdef MFS_SHIM
#include "mfs_gpu_shim.h"
#endif

// Intermediate globals defined in para_field_global.cpp but not declared in
// its header.
extern RealFlow tref, sref, amuref, trat, alpha_t, beita_t, temp, p_stag_t, rho_stag;
extern RealFlow ratio_rhop_max, ratio_rhop_min, ratio_p_break;

static void set_flow(double mach_in, double alpha_in, double cfl) {
    // Same formulas and order as para_field_global.cpp:37-209.
    mach = mach_in;
    alpha = alpha_in;
    cfl_start = cfl;
    cfl_end = cfl;
    cfl_min = 0.5 * cfl_start;
    cfl_ratio = cfl_end / cfl_start;
    alf_l = epsa_r;
    alf_n = epsa_r;
    trat = (T / tref);
    amu = amuref * trat * sqrt(trat) * (tref + sref) / (T + sref);
    ainf = sqrt(gam * gascon * T);
    uqq = ainf * mach;
    alpha_t = (alpha * PI / 180.);
    beita_t = (beita * PI / 180.);
    u_s = uqq * cos(alpha_t) * cos(-1. * beita_t);
    v_s = uqq * sin(-1. * beita_t);
    w_s = uqq * sin(alpha_t) * cos(-1. * beita_t);
    p_s = 0.00;
    rho_s = re * amu / uqq;
    p_bar = rho_s * gascon * T;
    gamm1 = gam - 1.0;
    temp = 1.0 + 0.5 * gamm1 * mach * mach;
    p_stag_t = pow(temp, gam / gamm1);
    rho_stag = pow(temp, 1.0 / gamm1) * rho_s;
    p_stag = (p_stag_t - 1.0) * p_bar;
    e_stag = p_bar / gamm1 + 0.5 * rho_s * (u_s * u_s + v_s * v_s + w_s * w_s);
    p_min = ratio_rhop_min * p_bar - p_bar;
    p_max = ratio_rhop_max * (p_stag + p_bar) - p_bar;
    rho_min = ratio_rhop_min * rho_s;
    rho_max = ratio_rhop_max * rho_stag;
    p_break = ratio_p_break * p_bar - p_bar;
    e_stag_max = 1.0e5 * e_stag;
    iter_done = 0;
    dt_max = 0.0;
    dt_min = BIG;
}

static void dump(const std::string& dir, const std::string& name, const void* p, size_t bytes) {
    std::string path = dir + "/" + name;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    if (bytes && fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "short write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

static void dump_state(PolyGrid* g, const std::string& dir, int it) {
    const int nC = g->GetNTCell(), nB = g->GetNBFace(), n = nC + nB;
    char pre[64];
    snprintf(pre, sizeof(pre), "it%d_", it);
    std::string P(pre);
    dump(dir, P + "rho.f64", g->GetRho(), sizeof(double) * n);
    dump(dir, P + "u.f64", g->GetU(), sizeof(double) * n);
    dump(dir, P + "v.f64", g->GetV(), sizeof(double) * n);
    dump(dir, P + "w.f64", g->GetW(), sizeof(double) * n);
    dump(dir, P + "p.f64", g->GetP(), sizeof(double) * n);
    if (g->GetRes()) dump(dir, P + "res.f64", g->GetRes(), sizeof(double) * 5 * nC);
    if (g->GetDt()) dump(dir, P + "dt.f64", g->GetDt(), sizeof(double) * nC);
    if (g->GetDQ()) dump(dir, P + "dq.f64", g->GetDQ(), sizeof(double) * 5 * n);
    dump(dir, P + "dqdx.f64", g->GetDqdx(), sizeof(double) * 5 * n);
    dump(dir, P + "dqdy.f64", g->GetDqdy(), sizeof(double) * 5 * n);
    dump(dir, P + "dqdz.f64", g->GetDqdz(), sizeof(double) * 5 * n);
}

int main(int argc, char** argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: %s MESH BC OUTDIR MACH ALPHA CFL ITERS DUMPLIST [ghosts] [ramp=CFL_START,NSTEP]\n", argv[0]);
        return 2;
    }
    // "ghosts": call SetGhostVariables after PassInterfaceData in every
    // iteration, where src/main.cpp:283 has the call commented out.
    // "ramp=a,n": CFL ramp from cfl_start = a to cfl_end = CFL over n
    // iterations (compute_timestep.cpp:868-871), cfl_min = 0.5 cfl_start as
    // in para_field_global.cpp:58.
    bool ghost_refresh = false;
    bool dump_geometry = true;  // "nogeo": skip the set-up geometry dumps (large meshes)
    double ramp_start = 0.0;
    int ramp_nstep = 0;
    std::string init_dir;
    for (int a = 9; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "ghosts") ghost_refresh = true;
        if (s == "nogeo") dump_geometry = false;
        if (s.rfind("ramp=", 0) == 0) {
            ramp_start = atof(s.substr(5).c_str());
            ramp_nstep = atoi(s.substr(s.find(',') + 1).c_str());
        }
        if (s.rfind("limits=", 0) == 0) {
            if (sscanf(s.c_str() + 7, "%lf,%lf,%lf", &ratio_rhop_min, &ratio_rhop_max, &ratio_p_break) != 3) {
                fprintf(stderr, "bad %s\n", s.c_str());
                return 2;
            }
        }
        if (s.rfind("dtratio=", 0) == 0) ratio_dtmax = atof(s.c_str() + 8);
        if (s.rfind("init=", 0) == 0) init_dir = s.substr(5);
    }
    const std::string mesh = argv[1], bcf = argv[2], out = argv[3];
    const double m_in = atof(argv[4]), a_in = atof(argv[5]), cfl = atof(argv[6]);
    const int iters = atoi(argv[7]);
    std::set<int> dumps;
    {
        std::string s = argv[8];
        size_t p = 0;
        while (p < s.size()) {
            size_t q = s.find(',', p);
            if (q == std::string::npos) q = s.size();
            if (q > p) dumps.insert(atoi(s.substr(p, q - p).c_str()));
            p = q + 1;
        }
    }
    set_flow(m_in, a_in, cfl);
    if (ramp_nstep > 0) {
        cfl_start = ramp_start;
        cfl_nstep = ramp_nstep;
        cfl_min = 0.5 * cfl_start;
        cfl_ratio = cfl_end / cfl_start;
    }
    n_steps = iters;

    PolyGrid* grid = new PolyGrid();
    BCond* bc = new BCond();
    ReadMFlowGrid_binary(grid, mesh);
    ReadMFlowBoundaryCondition(bc, bcf.c_str());
    Decoupling(grid);  // main.cpp:168; face colouring in the OpenMP build, empty in the serial one
    PreprocessGrid(grid, bc);
    if (!init_dir.empty()) {
        const size_t n = (size_t)grid->GetNTCell() + grid->GetNBFace();
        RealFlow* dst[5] = {grid->GetRho(), grid->GetU(), grid->GetV(), grid->GetW(), grid->GetP()};
        const char* nm[5] = {"rho", "u", "v", "w", "p"};
        for (int j = 0; j < 5; ++j) {
            const std::string path = init_dir + "/init_" + nm[j] + ".f64";
            FILE* f = fopen(path.c_str(), "rb");
            if (!f || fread(dst[j], sizeof(double), n, f) != n) {
                fprintf(stderr, "cannot read %zu doubles from %s\n", n, path.c_str());
                return 2;
            }
            fclose(f);
        }
#ifdef MFS_SHIM
        MfsUploadState(grid);
#endif
    }
#ifdef MFS_SHIM
    if (ghost_refresh) {
        fprintf(stderr, "ghosts: the shim driver keeps SetGhostVariables on the host; not supported\n");
        return 2;
    }
    MfsDownloadResults(grid);
#endif

    const int nC = grid->GetNTCell(), nB = grid->GetNBFace(), nF = grid->GetNTFace(), nN = grid->GetNTNode();
    if (dump_geometry) {
        const int n = nC + nB;
        dump(out, "geo_xfn.f64", grid->GetXfn(), sizeof(double) * nF);
        dump(out, "geo_yfn.f64", grid->GetYfn(), sizeof(double) * nF);
        dump(out, "geo_zfn.f64", grid->GetZfn(), sizeof(double) * nF);
        dump(out, "geo_area.f64", grid->GetFaceArea(), sizeof(double) * nF);
        dump(out, "geo_xfc.f64", grid->GetXfc(), sizeof(double) * nF);
        dump(out, "geo_yfc.f64", grid->GetYfc(), sizeof(double) * nF);
        dump(out, "geo_zfc.f64", grid->GetZfc(), sizeof(double) * nF);
        dump(out, "geo_xcc.f64", grid->GetXcc(), sizeof(double) * n);
        dump(out, "geo_ycc.f64", grid->GetYcc(), sizeof(double) * n);
        dump(out, "geo_zcc.f64", grid->GetZcc(), sizeof(double) * n);
        dump(out, "geo_vol.f64", grid->GetCellVol(), sizeof(double) * n);
        dump(out, "geo_wnode.f64", grid->GetWeightNodeDist(), sizeof(double) * nN);
        dump(out, "geo_nmark.i32", grid->GetNodeType(), sizeof(int) * nN);
        dump(out, "geo_f2c.i32", grid->Getf2c(), sizeof(int) * 2 * nF);
    }
    if (dumps.count(0)) dump_state(grid, out, 0);

    // history: per iteration norm[5], dt_max, dt_min
    double* hist = (double*)calloc((size_t)iters * 7 + 1, sizeof(double));
    double t_loop = 0.0;
    for (int it = 0; it < iters; ++it) {
        auto t0 = std::chrono::steady_clock::now();
        ZeroResiduals(grid);
        InviscidFlux(grid, NULL, 0);
        ComputeTimeStep(grid);
        ForwardLUSGS(grid, 0);
        PassInterfaceData(grid);
        if (ghost_refresh) SetGhostVariables(grid);
        CalculateGradient(grid);
        iter_done++;
#ifdef MFS_SHIM
        MfsDownloadResults(grid);  // the history below reads res and dt_max/dt_min on the host
#endif
        auto t1 = std::chrono::steady_clock::now();
        t_loop += std::chrono::duration<double>(t1 - t0).count();
        const double* res = grid->GetRes();
        for (int j = 0; j < 5; ++j) {
            double s = 0.0;
            for (int i = 0; i < nC; ++i) s += res[(size_t)j * nC + i] * res[(size_t)j * nC + i];
            hist[7 * it + j] = sqrt(s);
        }
        hist[7 * it + 5] = dt_max;
        hist[7 * it + 6] = dt_min;
        if (dumps.count(it + 1)) dump_state(grid, out, it + 1);
    }
    dump(out, "history.f64", hist, sizeof(double) * 7 * iters);
    FILE* f = fopen((out + "/summary.txt").c_str(), "w");
    if (f) {
        fprintf(f, "nC %d nF %d nB %d nN %d iters %d t_loop %.6f mcellit_s %.6f\n", nC, nF, nB, nN, iters, t_loop,
                t_loop > 0 ? (double)nC * iters / t_loop / 1e6 : 0.0);
        fprintf(f, "derived u_s %.17g w_s %.17g rho_s %.17g p_bar %.17g p_min %.17g p_max %.17g rho_min %.17g rho_max %.17g p_break %.17g\n",
                u_s, w_s, rho_s, p_bar, p_min, p_max, rho_min, rho_max, p_break);
        if (iters > 0)
            fprintf(f, "%5d %.5e %.5e %.5e %.5e %.5e %.5e %.5e\n", iters, hist[7 * (iters - 1)], hist[7 * (iters - 1) + 1],
                    hist[7 * (iters - 1) + 2], hist[7 * (iters - 1) + 3], hist[7 * (iters - 1) + 4], dt_max, dt_min);
        fclose(f);
    }
    free(hist);
    printf("ref_driver done: nC=%d iters=%d t_loop=%.3fs (%.3f M cell-iter/s)\n", nC, iters, t_loop,
           t_loop > 0 ? (double)nC * iters / t_loop / 1e6 : 0.0);
    return 0;
}
