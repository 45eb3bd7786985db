// C ABI: parameters, errors and the host grid (LoadGrid / Decoupling /
// PreprocessGrid equivalents).  Pure host code.
#include <cstring>
#include <string>

#include "../../include/mfs.h"
#include "mfs_grid.hpp"
#include "mfs_mesh.hpp"
#include "mfs_tdtree.hpp"

namespace mfs {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace mfs

struct mfs_grid {
    mfs::Grid g;
};

using namespace mfs;

extern "C" const char* mfs_last_error(void) { return g_last_error.c_str(); }

extern "C" void mfs_params_default(mfs_params* p) {
    if (p) params_default(*p);
}

extern "C" int mfs_params_derive(mfs_params* p) {
    if (!p) {
        set_error("mfs_params_derive: null");
        return MFS_ERR_ARG;
    }
    if (!(p->mach > 0.0) || !(p->T > 0.0) || !(p->re > 0.0) || !(p->gam > 1.0) || !(p->cfl_start > 0.0) ||
        p->cfl_nstep <= 0) {
        set_error("mfs_params_derive: mach, T, re, cfl_start must be positive, gam > 1, cfl_nstep > 0");
        return MFS_ERR_ARG;
    }
    params_derive(*p);
    return MFS_OK;
}

extern "C" int mfs_grid_generate_box_bump(int nx, int ny, int nz, double jitter, double bump_height, double bump_sigma,
                                          uint64_t seed, mfs_grid** out) {
    return mfs_grid_generate_box(nx, ny, nz, MFS_CELLS_TET, jitter, bump_height, bump_sigma, seed, out);
}

extern "C" int mfs_grid_generate_box(int nx, int ny, int nz, int cells, double jitter, double bump_height,
                                     double bump_sigma, uint64_t seed, mfs_grid** out) {
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(jitter >= 0.0 && jitter < 0.25) || !(bump_sigma > 0.0) ||
        (cells != MFS_CELLS_TET && cells != MFS_CELLS_HEX && cells != MFS_CELLS_PRISM)) {
        set_error("mfs_grid_generate_box: bad argument (n >= 1, 0 <= jitter < 0.25, sigma > 0, known cell kind)");
        return MFS_ERR_ARG;
    }
    if ((double)nx * ny * nz * 6.0 > 2.0e9) {
        set_error("mfs_grid_generate_box: mesh too large for 32-bit indices");
        return MFS_ERR_ARG;
    }
    BoxBumpSpec s;
    s.cells = cells;
    s.nx = nx;
    s.ny = ny;
    s.nz = nz;
    s.jitter = jitter;
    s.bump_height = bump_height;
    s.bump_sigma = bump_sigma;
    s.seed = seed;
    mfs_grid* g = new mfs_grid();
    g->g.m = make_box_bump(s);
    *out = g;
    return MFS_OK;
}

extern "C" int mfs_grid_load(const char* mesh_path, const char* bc_path, mfs_grid** out) {
    if (!mesh_path || !bc_path || !out) {
        set_error("mfs_grid_load: null argument");
        return MFS_ERR_ARG;
    }
    mfs_grid* g = new mfs_grid();
    std::string err;
    if (!read_flowstar(g->g.m, mesh_path, bc_path, &err)) {
        delete g;
        set_error("mfs_grid_load: " + err);
        return MFS_ERR_IO;
    }
    *out = g;
    return MFS_OK;
}

extern "C" int mfs_grid_save(const mfs_grid* g, const char* mesh_path, const char* bc_path) {
    if (!g || !mesh_path || !bc_path) {
        set_error("mfs_grid_save: null argument");
        return MFS_ERR_ARG;
    }
    if (g->g.preprocessed) {
        set_error("mfs_grid_save: save the grid before PreprocessGrid (SpecifyBC rewrites f2c)");
        return MFS_ERR_STATE;
    }
    std::string err;
    if (!write_flowstar(g->g.m, mesh_path, bc_path, &err)) {
        set_error("mfs_grid_save: " + err);
        return MFS_ERR_IO;
    }
    return MFS_OK;
}

extern "C" int mfs_grid_decouple(mfs_grid* g) { return mfs_grid_decouple_ex(g, 4, 512); }

extern "C" int mfs_grid_decouple_ex(mfs_grid* g, int nb_parts, int part_size) {
    if (!g) {
        set_error("mfs_grid_decouple: null");
        return MFS_ERR_ARG;
    }
    if (nb_parts < 2 || nb_parts > 4 || part_size < 8) {
        set_error("mfs_grid_decouple_ex: need 2 <= nb_parts <= 4 and part_size >= 8");
        return MFS_ERR_ARG;
    }
    std::string e = g->g.decouple(nb_parts, part_size);
    if (!e.empty()) {
        set_error("mfs_grid_decouple: " + e);
        return MFS_ERR_STATE;
    }
    return MFS_OK;
}

extern "C" int mfs_grid_preprocess(mfs_grid* g, const mfs_params* p) {
    if (!g || !p) {
        set_error("mfs_grid_preprocess: null argument");
        return MFS_ERR_ARG;
    }
    std::string e = g->g.preprocess(*p);
    if (!e.empty()) {
        set_error("mfs_grid_preprocess: " + e);
        return MFS_ERR_MESH;
    }
    return MFS_OK;
}

extern "C" int mfs_grid_preprocess_device(mfs_grid* g, const mfs_params* p) {
    if (!g || !p) {
        set_error("mfs_grid_preprocess_device: null argument");
        return MFS_ERR_ARG;
    }
    std::string e = g->g.preprocess(*p, true);
    if (!e.empty()) {
        set_error("mfs_grid_preprocess_device: " + e);
        return MFS_ERR_MESH;
    }
    return MFS_OK;
}

extern "C" int mfs_grid_get_info(const mfs_grid* g, mfs_grid_info* info) {
    if (!g || !info) {
        set_error("mfs_grid_get_info: null argument");
        return MFS_ERR_ARG;
    }
    info->nTCell = g->g.m.nC;
    info->nTFace = g->g.m.nF;
    info->nBFace = g->g.m.nB;
    info->nTNode = g->g.m.nN;
    info->preprocessed = g->g.preprocessed ? 1 : 0;
    info->decoupled = g->g.decoupled ? 1 : 0;
    return MFS_OK;
}

namespace {
struct ArrayRef {
    const void* p = nullptr;
    size_t bytes = 0;
};

template <typename T>
ArrayRef ref(const std::vector<T>& v) {
    return {v.data(), v.size() * sizeof(T)};
}

ArrayRef find_array(const mfs_grid* gg, const char* name, std::vector<int>* scratch) {
    const Grid& g = gg->g;
    const std::string s = name;
    if (s == "x") return ref(g.m.x);
    if (s == "y") return ref(g.m.y);
    if (s == "z") return ref(g.m.z);
    if (s == "nNPF") return ref(g.m.nNPF);
    if (s == "f2n") return ref(g.m.f2n);
    if (s == "f2c") return ref(g.m.f2c);
    if (s == "bctype") return ref(g.bctype);
    if (s == "xfn") return ref(g.xfn);
    if (s == "yfn") return ref(g.yfn);
    if (s == "zfn") return ref(g.zfn);
    if (s == "area") return ref(g.area);
    if (s == "xfc") return ref(g.xfc);
    if (s == "yfc") return ref(g.yfc);
    if (s == "zfc") return ref(g.zfc);
    if (s == "xcc") return ref(g.xcc);
    if (s == "ycc") return ref(g.ycc);
    if (s == "zcc") return ref(g.zcc);
    if (s == "vol") return ref(g.vol);
    if (s == "rho") return ref(g.rho);
    if (s == "u") return ref(g.u);
    if (s == "v") return ref(g.v);
    if (s == "w") return ref(g.w);
    if (s == "p") return ref(g.p);
    if (s == "cellPerm") return ref(g.tree.cellPerm);
    if (s == "facePerm") return ref(g.tree.facePerm);
    if (s == "nodePerm") return ref(g.tree.nodePerm);
    if (s == "tdtree_nodes" && scratch) {
        // per tree node: firstCell lastCell firstFace lastFace firstNode lastNode isLeaf isIso nbParts iso son0..son3
        scratch->clear();
        for (const TDTreeNode& d : g.tree.nodes) {
            int rec[14] = {d.firstCell, d.lastCell, d.firstFace, d.lastFace, d.firstNode, d.lastNode,
                           d.isLeaf ? 1 : 0, d.isIso ? 1 : 0, d.nbParts, d.iso, -1, -1, -1, -1};
            for (size_t k = 0; k < d.son.size() && k < 4; ++k) rec[10 + k] = d.son[k];
            scratch->insert(scratch->end(), rec, rec + 14);
        }
        return {scratch->data(), scratch->size() * sizeof(int)};
    }
    return {};
}
}  // namespace

extern "C" size_t mfs_grid_array_bytes(const mfs_grid* g, const char* name) {
    if (!g || !name) return 0;
    std::vector<int> scratch;
    return find_array(g, name, &scratch).bytes;
}

extern "C" int mfs_grid_get_array(const mfs_grid* g, const char* name, void* dst, size_t bytes) {
    if (!g || !name || !dst) {
        set_error("mfs_grid_get_array: null argument");
        return MFS_ERR_ARG;
    }
    std::vector<int> scratch;
    ArrayRef a = find_array(g, name, &scratch);
    if (!a.p && a.bytes == 0) {
        set_error(std::string("mfs_grid_get_array: unknown or empty array ") + name);
        return MFS_ERR_ARG;
    }
    if (a.bytes != bytes) {
        set_error(std::string("mfs_grid_get_array: size mismatch for ") + name);
        return MFS_ERR_ARG;
    }
    std::memcpy(dst, a.p, bytes);
    return MFS_OK;
}

extern "C" void mfs_grid_destroy(mfs_grid* g) { delete g; }

extern "C" int mfs_grid_view(const mfs_grid* gg, mfs_mesh_view* v) {
    if (!gg || !v) {
        set_error("mfs_grid_view: null argument");
        return MFS_ERR_ARG;
    }
    const Grid& g = gg->g;
    if (!g.preprocessed) {
        set_error("mfs_grid_view: grid is not preprocessed");
        return MFS_ERR_STATE;
    }
    v->nTCell = g.m.nC;
    v->nTFace = g.m.nF;
    v->nBFace = g.m.nB;
    v->nTNode = g.m.nN;
    v->x = g.m.x.data();
    v->y = g.m.y.data();
    v->z = g.m.z.data();
    v->nNPF = g.m.nNPF.data();
    v->f2n = g.m.f2n.data();
    v->f2c = g.m.f2c.data();
    v->bc_type = g.bctype.data();
    v->xfn = g.xfn.data();
    v->yfn = g.yfn.data();
    v->zfn = g.zfn.data();
    v->area = g.area.data();
    v->xfc = g.xfc.data();
    v->yfc = g.yfc.data();
    v->zfc = g.zfc.data();
    v->xcc = g.xcc.data();
    v->ycc = g.ycc.data();
    v->zcc = g.zcc.data();
    v->vol = g.vol.data();
    v->rho = g.rho.data();
    v->u = g.u.data();
    v->v = g.v.data();
    v->w = g.w.data();
    v->p = g.p.data();
    return MFS_OK;
}
