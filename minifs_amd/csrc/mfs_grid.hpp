// Host grid of the solver path: the PolyGrid state after the set-up of
// src/main.cpp:166-170 (LoadGrid, Decoupling, PreprocessGrid).
#pragma once
#include <string>
#include <vector>

#include "../../include/mfs.h"
#include "mfs_mesh.hpp"
#include "mfs_tdtree.hpp"

namespace mfs {

// BIG / TINY of src/Base/number_type.h:28-29.
constexpr double kBIG = 1.e40;
constexpr double kTINY = 1.e-40;
constexpr double kPI = 3.14159265358979323846;

struct Grid {
    RawMesh m;
    bool decoupled = false;
    bool preprocessed = false;
    TDTree tree;
    mfs_params params{};
    std::vector<int> bctype;                            // nB, from the bc.para record of each boundary face
    std::vector<double> xfn, yfn, zfn, area, xfc, yfc, zfc;  // nF
    std::vector<double> xcc, ycc, zcc, vol;             // nC + nB
    std::vector<double> rho, u, v, w, p;                // nC + nB

    // Set-up steps; return "" on success or an error message.
    std::string decouple(int nbParts = 4, int partSize = 512);
    std::string preprocess(const mfs_params& prm, bool device_metrics = false);
};

// Pieces of PreprocessGrid, exposed for tests.
std::string specify_bc(Grid& g);
std::string compute_metrics(Grid& g);
// The scale of CorrectCellCentroid's push, 10^(1-16), shared by both paths.
double centroid_scale();
// The same on the current HIP device (mfs_device.hip), bitwise equal.
std::string compute_metrics_device(Grid& g);
void init_flowfield(Grid& g);
void set_ghost_variables(Grid& g);

void params_default(mfs_params& p);
void params_derive(mfs_params& p);

}  // namespace mfs
