// Host-side mesh model of the MI355X MiniFS solver path.
//
// RawMesh is the FlowStar binary mesh exactly as the reference reads it
// (src/Kernels/load_grid.cpp:129-273): node coordinates, face->node lists,
// face->cell pairs with boundary sides stored as -patch_id, boundary faces
// first.  The box+bump generator, the FlowStar reader/writer and the TDTree
// renumbering all operate on it.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mfs {

// Boundary-condition codes (src/Base/constant.h:48-52).
enum BcType : int { BC_WALL = 3, BC_SYMM = 4, BC_FAR_FIELD = 6, BC_FAR_FIELD_MOD = 61, BC_INTERFACE = 10 };

struct RawMesh {
    int nN = 0, nF = 0, nC = 0;
    int nB = 0;                    // counted by the reader (load_grid.cpp:206-232)
    std::vector<double> x, y, z;   // nN
    std::vector<int> nNPF;         // nF
    std::vector<int> f2n;          // sum(nNPF), concatenated
    std::vector<int> f2c;          // 2*nF; second entry of a boundary face = -patch_id
    // bc.para records in file order: record k has patch id patch[k] and type type[k];
    // a boundary face with side -p uses record p-1 (load_grid.cpp:883-901).
    std::vector<int> bc_patch, bc_type;

    std::vector<int> f2n_offsets() const;
};

// Cell kinds of the box generator.
enum BoxCells : int { CELLS_TET = 0, CELLS_HEX = 1, CELLS_PRISM = 2 };

struct BoxBumpSpec {
    int nx = 20, ny = 20, nz = 20;
    int cells = CELLS_TET;       // tetrahedra (6 per hex), hexahedra, or triangular prisms (2 per hex)
    double jitter = 0.10;        // interior node jitter, fraction of the local spacing
    double bump_height = 0.05;   // Gaussian bump on the z=0 wall
    double bump_sigma = 0.10;
    uint64_t seed = 20260101ULL;
};

// Unit box, nx*ny*nz hexes, each split into 6 Kuhn tetrahedra around the
// (0,0,0)-(1,1,1) diagonal; interior nodes jittered; z=0 wall (patch 1, WALL)
// displaced by the bump, blended linearly to zero at z=1; other five sides
// are patch 2 (FAR_FIELD).  Boundary faces first, faces oriented so the
// right-hand normal points from c1 to c2 (outward on the boundary).
// With cells = CELLS_HEX each hex is one cell (6 quadrilateral faces); with
// CELLS_PRISM each hex is split along its (0,0)-(1,1) xy diagonal into two
// triangular prisms (2 triangles + 3 quadrilaterals).  Same nodes, bump,
// patches and face ordering as the tetrahedral mesh.
RawMesh make_box_bump(const BoxBumpSpec& spec);

// FlowStar binary writer/reader (format of load_grid.cpp:129-273, with the
// trailing nIFace=0,nINode=0 record the reader requires) and bc.para text.
bool write_flowstar(const RawMesh& m, const std::string& mesh_path, const std::string& bc_path, std::string* err);
bool read_flowstar(RawMesh& m, const std::string& mesh_path, const std::string& bc_path, std::string* err);

// Splitmix64-based uniform in [-1,1) used by the generator.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform_pm1() { return (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }
    double uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace mfs
