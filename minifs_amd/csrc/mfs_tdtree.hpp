// TDTree recursive decomposition and renumbering of a RawMesh.
//
// Restates the tree construction of src/src/TDTree_creation.cpp:407-519
// (root -> physical-boundary face tree, interface face tree, recursive cell
// tree with 4 parts + separator, leaves below 512 items) and the
// renumbering of src/Base/grid_polyhedra.cpp:88-353 (cells, nodes, faces).
// The graph partitioner is this project's own recursive coordinate
// bisection of the nodal graph's vertices; the reference calls METIS 5.1
// METIS_PartGraphRecursive (src/src/TDTree_partitioning.cpp:79-81) there.
#pragma once
#include <vector>

#include "mfs_mesh.hpp"

namespace mfs {

struct TDTreeNode {
    int firstCell = 0, lastCell = -1;
    int firstFace = 0, lastFace = -1;
    int firstNode = 0, lastNode = -1;
    bool isIso = false, isLeaf = true;
    int nbParts = 0;
    std::vector<int> son;  // indices into TDTree::nodes, -1 for an empty part
    int iso = -1;
};

struct TDTree {
    std::vector<TDTreeNode> nodes;  // nodes[0] is the root
    // perm[old] = new position; rev[new] = old
    std::vector<int> cellPerm, cellRev, facePerm, faceRev, nodePerm, nodeRev;
    int partSize = 512;
    int nbParts = 4;

    int root() const { return 0; }
    // Cell-tree leaves in forward traversal order (sons before the separator).
    void leaves_forward(int node, std::vector<int>& out) const;
};

// Builds the tree on the mesh in its current numbering.
TDTree build_tdtree(const RawMesh& m, int nbParts = 4, int partSize = 512);

// Applies the tree's permutations (TDTree_CellReordering, _NodeReordering,
// _FaceReordering) and returns the renumbered mesh.
RawMesh renumber(const RawMesh& m, const TDTree& t);

}  // namespace mfs
