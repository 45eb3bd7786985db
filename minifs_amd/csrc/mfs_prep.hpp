// Host-side construction of the device layout from a preprocessed mesh view.
//
// Every per-cell or per-node sum of the reference is a face-ordered (or
// cell-ordered) serial accumulation.  The device kernels gather instead of
// scatter, so this layout stores, for each cell and each node, its
// contributions in exactly the order the reference's serial loops add them;
// a gather that walks these lists in order performs the same floating-point
// additions in the same order as the reference's scatter.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mfs.h"

namespace mfs {

struct HostLayout {
    int nC = 0, nF = 0, nB = 0, nN = 0;
    // faces
    std::vector<int> c1, c2;                   // nF (c2 = nC+k for boundary face k)
    std::vector<double> fgeo;                  // 4*nF: xfn, yfn, zfn, area
    std::vector<double> dn1, dn2;              // nF: |(x_fc - x_cc) . n| for each side
    std::vector<int> fno, fn;                  // face->node CSR (F2N)
    std::vector<int> btype;                    // nB
    // cells: faces in ascending face id (CalC2F, grid_polyhedra_auxiliary_functions.cpp:1345-1399)
    std::vector<int> cfo, cf;                  // nC+1, sum nFPC
    std::vector<double> vol;                   // nC
    // node values (CompNodeVar3D_dist): per node, ordered entries
    // value = q[a] (b < 0) or 0.5*(q[a]+q[b]) (b >= 0, boundary face average)
    std::vector<int> nlo;                      // nN+1
    std::vector<int> nla, nlb;                 // entries
    std::vector<double> nlw;                   // entry weights
    std::vector<double> nden;                  // nN: WeightNode + TINY
    std::vector<int> nmark;                    // nN (0, WALL, FAR_FIELD, 101)
    std::vector<int> sym_node;                 // nodes with nodesymm == 1
    std::vector<double> sym_n;                 // 3 per symmetry node (xfn_n_symm, ...)
    // LU-SGS sweep schedule
    std::vector<int> flevel, blevel;           // nC: forward / backward DAG level
    std::vector<int> sched_f, sched_b;         // cells sorted by (level, id)
    int nlev_f = 0, nlev_b = 0;
    // initial state, 5 blocks of nC+nB
    std::vector<double> q;
};

// Returns "" on success.
std::string build_layout(const mfs_mesh_view& v, HostLayout& L);

}  // namespace mfs
