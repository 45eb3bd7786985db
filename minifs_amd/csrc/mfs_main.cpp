// mfs_solve: the solver's native driver, the counterpart of the reference's
// main program (src/main.cpp:64-345).  It builds the grid (a FlowStar mesh
// with its bc.para, or the box+bump generator), renumbers it (TDTree
// decoupling), preprocesses it, runs the iteration loop of main.cpp:198-291
// on the GPU and prints the reference's residual line (main.cpp:294-329)
// and the per-kernel times.
//
//   mfs_solve [--mesh FILE --bc FILE | --box N [--cells tet|hex|prism]]
//             [--mach M] [--alpha DEG] [--cfl C] [--steps K] [--warmup W]
//             [--precision fp64|mixed] [--entry-points] [--timing] [--json]
//
// --entry-points runs each iteration through the reference's kernel entry
// points (ZeroResiduals, InviscidFlux, ComputeTimeStep, ForwardLUSGS,
// PassInterfaceData, CalculateGradient) instead of the fused mfs_step.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mfs.h"

namespace {

struct Args {
    std::string mesh, bc;
    int box = 40;
    int cells = MFS_CELLS_TET;
    double mach = 0.5, alpha = 0.0, cfl = 10.0;
    int steps = 20, warmup = 3;
    int precision = MFS_FP64;
    bool entry_points = false, timing = false, json = false;
    int tree_parts = 4, tree_leaf = 512;  // TDTree parameters (CreateTree's 4 and 512)
    bool device_metrics = false;          // ComputeMetrics on the GPU
};

int usage() {
    std::fprintf(stderr,
                 "usage: mfs_solve [--mesh FILE --bc FILE | --box N [--cells tet|hex|prism]] [--mach M] "
                 "[--alpha DEG] [--cfl C] [--steps K] [--warmup W] [--precision fp64|mixed] [--entry-points] "
                 "[--timing] [--json] [--tree PARTS,LEAF] [--device-metrics]\n");
    return 2;
}

bool parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&](const char*& v) {
            if (i + 1 >= argc) return false;
            v = argv[++i];
            return true;
        };
        const char* v = nullptr;
        if (k == "--entry-points") {
            a.entry_points = true;
        } else if (k == "--timing") {
            a.timing = true;
        } else if (k == "--json") {
            a.json = true;
        } else if (k == "--device-metrics") {
            a.device_metrics = true;
        } else if (!val(v)) {
            return false;
        } else if (k == "--mesh") {
            a.mesh = v;
        } else if (k == "--bc") {
            a.bc = v;
        } else if (k == "--box") {
            a.box = std::atoi(v);
        } else if (k == "--cells") {
            const std::string c = v;
            if (c == "tet") a.cells = MFS_CELLS_TET;
            else if (c == "hex") a.cells = MFS_CELLS_HEX;
            else if (c == "prism") a.cells = MFS_CELLS_PRISM;
            else return false;
        } else if (k == "--mach") {
            a.mach = std::atof(v);
        } else if (k == "--alpha") {
            a.alpha = std::atof(v);
        } else if (k == "--cfl") {
            a.cfl = std::atof(v);
        } else if (k == "--steps") {
            a.steps = std::atoi(v);
        } else if (k == "--warmup") {
            a.warmup = std::atoi(v);
        } else if (k == "--tree") {
            if (std::sscanf(v, "%d,%d", &a.tree_parts, &a.tree_leaf) != 2) return false;
        } else if (k == "--precision") {
            const std::string c = v;
            if (c == "fp64") a.precision = MFS_FP64;
            else if (c == "mixed") a.precision = MFS_MIXED;
            else return false;
        } else {
            return false;
        }
    }
    return a.box > 0 && a.steps >= 0 && a.warmup >= 0 && (a.mesh.empty() == a.bc.empty());
}

int fail(const char* what) {
    std::fprintf(stderr, "mfs_solve: %s: %s\n", what, mfs_last_error());
    return 1;
}

// one iteration of main.cpp:198-291 through the kernel entry points
int iterate_entry_points(mfs_ctx* c) {
    int rc = mfs_zero_residuals(c);
    if (rc == MFS_OK) rc = mfs_inviscid_flux(c);
    if (rc == MFS_OK) rc = mfs_compute_timestep(c);
    if (rc == MFS_OK) rc = mfs_forward_lusgs(c);
    if (rc == MFS_OK) rc = mfs_pass_interface_data(c);
    if (rc == MFS_OK) rc = mfs_calculate_gradient(c);
    if (rc == MFS_OK) rc = mfs_iteration_done(c);
    return rc;
}

int run(mfs_ctx* c, const Args& a, int n) {
    if (!a.entry_points) return mfs_step(c, n);
    for (int i = 0; i < n; ++i) {
        const int rc = iterate_entry_points(c);
        if (rc != MFS_OK) return rc;
    }
    return MFS_OK;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a)) return usage();

    mfs_params p;
    mfs_params_default(&p);
    p.mach = a.mach;
    p.alpha = a.alpha;
    p.cfl_start = p.cfl_end = a.cfl;
    if (mfs_params_derive(&p) != MFS_OK) return fail("parameters");

    const auto t0 = std::chrono::steady_clock::now();
    mfs_grid* g = nullptr;
    if (!a.mesh.empty()) {
        if (mfs_grid_load(a.mesh.c_str(), a.bc.c_str(), &g) != MFS_OK) return fail("LoadGrid");
    } else if (mfs_grid_generate_box(a.box, a.box, a.box, a.cells, 0.1, 0.05, 0.1, 20260101ULL, &g) != MFS_OK) {
        return fail("box mesh");
    }
    if (mfs_grid_decouple_ex(g, a.tree_parts, a.tree_leaf) != MFS_OK) return fail("Decoupling");
    if ((a.device_metrics ? mfs_grid_preprocess_device(g, &p) : mfs_grid_preprocess(g, &p)) != MFS_OK)
        return fail("PreprocessGrid");
    mfs_grid_info gi;
    mfs_mesh_view view;
    if (mfs_grid_get_info(g, &gi) != MFS_OK || mfs_grid_view(g, &view) != MFS_OK) return fail("grid view");
    mfs_ctx* c = nullptr;
    if (mfs_ctx_create(&view, &p, a.precision, &c) != MFS_OK) return fail("device context");
    const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (run(c, a, a.warmup) != MFS_OK || mfs_sync(c) != MFS_OK) return fail("warm-up");
    if (a.timing && (mfs_ctx_enable_timing(c, 1) != MFS_OK || mfs_ctx_reset_timing(c) != MFS_OK))
        return fail("timing");
    const auto t1 = std::chrono::steady_clock::now();
    if (run(c, a, a.steps) != MFS_OK || mfs_sync(c) != MFS_OK) return fail("iterations");
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();

    const int nh = mfs_history_count(c);
    std::vector<double> h(7 * (size_t)(nh > 0 ? nh : 1), 0.0);
    if (nh > 0 && mfs_get_history(c, h.data(), nh) != nh) return fail("history");
    const double* last = h.data() + 7 * (size_t)(nh > 0 ? nh - 1 : 0);

    static const char* kNames[] = {"flux", "timestep", "flux_ts", "reduce", "diag",
                                   "clamp", "fwd_sweep", "bwd_sweep", "node", "gradient", "grad_flux_ts", "ghosts"};
    if (a.json) {
        std::printf("{\"cells\": %d, \"faces\": %d, \"bfaces\": %d, \"nodes\": %d, \"precision\": \"%s\", "
                    "\"steps\": %d, \"ms_per_step\": %.4f, \"M_cell_iters_per_s\": %.3f, \"setup_s\": %.2f, "
                    "\"last_residual\": [%.17g, %.17g, %.17g, %.17g, %.17g], \"dt_max\": %.17g, \"dt_min\": %.17g",
                    gi.nTCell, gi.nTFace, gi.nBFace, gi.nTNode, a.precision == MFS_MIXED ? "mixed" : "fp64", a.steps,
                    a.steps ? el / a.steps * 1e3 : 0.0, a.steps && el > 0 ? gi.nTCell * (double)a.steps / el / 1e6 : 0.0,
                    setup_s, last[0], last[1], last[2], last[3], last[4], last[5], last[6]);
        if (a.timing) {
            std::printf(", \"kernel_ms_per_launch\": {");
            bool first = true;
            for (const char* k : kNames) {
                double ms = 0.0;
                int n = 0;
                if (mfs_ctx_get_timing(c, k, &ms, &n) == MFS_OK && n > 0) {
                    std::printf("%s\"%s\": %.4f", first ? "" : ", ", k, ms / n);
                    first = false;
                }
            }
            std::printf("}");
        }
        std::printf("}\n");
    } else {
        // main.cpp:320-324
        std::printf("\n#iter   rho_res      mx_res      my_res      mz_res     et_res     dt_max    dt_min\n");
        std::printf("%5d %.5e %.5e %.5e %.5e %.5e %.5e %.5e\n", a.warmup + a.steps, last[0], last[1], last[2],
                    last[3], last[4], last[5], last[6]);
        std::printf("cells %d, %d iterations in %.3f s: %.4f ms/iteration, %.1f M cell-iterations/s\n", gi.nTCell,
                    a.steps, el, a.steps ? el / a.steps * 1e3 : 0.0,
                    a.steps && el > 0 ? gi.nTCell * (double)a.steps / el / 1e6 : 0.0);
        if (a.timing)
            for (const char* k : kNames) {
                double ms = 0.0;
                int n = 0;
                if (mfs_ctx_get_timing(c, k, &ms, &n) == MFS_OK && n > 0)
                    std::printf("  %-10s %8.1f us/launch  (%d launches)\n", k, ms / n * 1e3, n);
            }
    }
    std::fflush(stdout);
    mfs_ctx_destroy(c);
    mfs_grid_destroy(g);
    return 0;
}
