// Host-side construction of the device layout from a preprocessed mesh view.
//
// Every per-cell or per-node sum of the reference is a face-ordered (or
// cell-ordered) serial accumulation.  The device kernels gather instead of
// scatter, so this layout stores, for each cell and each node, its
// contributions in exactly the order the reference's serial loops add them;
// a gather that walks these lists in order performs the same floating-point
// additions in the same order as the reference's scatter.
//
// Storage order on the device is free as long as that order of operations
// is kept.  Cells are stored in the order of the forward LU-SGS schedule
// (DAG level, then mesh id), so the cells of a sweep chunk are contiguous;
// faces and nodes are stored in first-touch order of that cell traversal.  Everything that depends on mesh numbering (per-cell face
// order, the sweep direction test "neighbour id < cell id", the per-node
// summation order) is evaluated here on mesh ids and stored as lists or
// masks.
#pragma once
#include <cstdint>
#include <string>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/mfs.h"

namespace mfs {

// Allocator whose resize() leaves new elements default-initialised (no
// zero fill): the layout's large arrays are then first touched by the
// threads that fill them, not by one thread's memset.
template <typename T>
struct NoFillAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoFillAlloc<U>;
    };
    NoFillAlloc() = default;
    template <typename U>
    NoFillAlloc(const NoFillAlloc<U>&) noexcept {}
    template <typename U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <typename U, typename... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
template <typename T>
using lvec = std::vector<T, NoFillAlloc<T>>;

struct HostLayout {
    int nC = 0, nF = 0, nB = 0, nN = 0;
    // cells in device order; cell_of_dev[d] = mesh cell id (ghost k stays at nC+k)
    lvec<int> cell_of_dev;
    int dev_cell0 = 0;                         // device id of mesh cell 0
    // per device cell: bits 0-7 forward dependency slots (interior neighbour
    // with a smaller mesh id), bits 8-15 backward slots (neighbour with a
    // larger mesh id, ghosts included), bits 16-23 backward slots whose
    // neighbour is an interior cell, bits 24-31 boundary slots whose face
    // value in the gradient is the cell/ghost average (not SYMM/INTERFACE)
    lvec<int> depmask;
    // faces, in device order; face_of_dev[d] = mesh face id
    lvec<int> face_of_dev;
    lvec<int> c1, c2;                   // nF, device cell ids (c2 = nC+k for boundary face k)
    lvec<double> fgeo;                  // 4*nF: xfn, yfn, zfn, area
    lvec<double> dn1, dn2;              // nF: |(x_fc - x_cc) . n| for each side
    lvec<int> fno, fn;                  // face->node CSR in device node ids
    lvec<int> btype;                    // nB, by boundary face (ghost) index k
    // cells: device face ids in ascending mesh face id (CalC2F,
    // grid_polyhedra_auxiliary_functions.cpp:1345-1399)
    lvec<int> cfo, cf;                  // nC+1, sum nFPC
    lvec<double> vol;                   // nC
    int max_faces = 0;
    // nodes, in device order; node_of_dev[d] = mesh node id
    lvec<int> node_of_dev;
    // node values (CompNodeVar3D_dist): per device node, ordered entries
    // value = q[a] (b < 0) or 0.5*(q[a]+q[b]) (b >= 0, boundary face average)
    lvec<int> nlo;                      // nN+1
    lvec<int> nla, nlb;                 // entries
    lvec<double> nlw;                   // entry weights
    lvec<double> nden;                  // nN: WeightNode + TINY
    lvec<int> sym_node;                 // device nodes with nodesymm == 1
    lvec<double> sym_n;                 // 3 per symmetry node (xfn_n_symm, ...)
    // LU-SGS sweeps: forward DAG level of each mesh cell, the number of
    // levels, and the chunks (device cell ranges of one level, `cpc` cells,
    // `lpc` lanes per cell) with their latest dependencies
    lvec<int> flevel;                   // nC, by mesh id
    int nlev = 0;
    int lpc = 4, cpc = 16;
    lvec<int> chunk;                    // 2 per chunk (device order): first device cell, cell count
    lvec<int> tile_chunk;               // ntile+1: first chunk of each storage tile
    lvec<int> run;                      // chunks by forward level (then tile): the flat schedule
    lvec<int> rep_f, rep_b;             // per chunk: dependency expected to be published last (-1: none)
    lvec<int> cfs;                      // lpc slots per device cell: device face id or -1
    // what the sweeps read, laid out so that a chunk needs no chain of
    // lookups: per position u of `run`, {first device cell, cell count
    // (| kThinChunkFlag on a thin level), rep_f, rep_b} of chunk run[u]; per
    // chunk, its position.  The per-slot tables (neighbour, oriented normal
    // and area, side distance, sides, tetrahedra's face nodes) are built on
    // the device from cfs and the face arrays (k_slot_tables).
    lvec<int> crun;                     // 4 per position
    lvec<int> pos;                      // per chunk
    bool tri_slots = false;                    // tetrahedra with triangular slots only
    // initial state, 5 blocks of nC+nB, device cell order
    lvec<double> q;
};

// Cells per storage tile of the device order (the sweeps' workgroup unit).
constexpr int kDefaultTileCells = 4096;
// Sweep levels of at most this many chunks are thin: their chunks are
// flagged in the count word of their record (crun[4u+1] bit 8).
constexpr int kThinLevelChunks = 64;
constexpr int kThinChunkFlag = 0x100;

// Returns "" on success.
std::string build_layout(const mfs_mesh_view& v, HostLayout& L, int tile_cells = kDefaultTileCells);

}  // namespace mfs
