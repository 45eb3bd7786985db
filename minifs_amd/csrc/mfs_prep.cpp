// Device-layout construction (see mfs_prep.hpp).
#include "mfs_prep.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <sstream>
#include <thread>

#include "mfs_mesh.hpp"

namespace mfs {


namespace {
constexpr double kTINY_ = 1.e-40;

// Runs fn(begin, end, t) over T contiguous ranges of [0, n) on T threads
// (the calling thread takes the last range).  Every result below is written
// by exactly one range and does not depend on T.
int layout_threads(size_t n) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t want = n / 65536 + 1;
    return (int)std::max<size_t>(1, std::min<size_t>({want, hw ? hw : 1, 16}));
}
template <typename F>
void parallel_ranges(size_t n, int T, F&& fn) {
    std::vector<std::thread> th;
    th.reserve(T > 0 ? T - 1 : 0);
    for (int t = 0; t + 1 < T; ++t) th.emplace_back([&, t] { fn(n * t / T, n * (t + 1) / T, t); });
    fn(n * (T - 1) / T, n, T - 1);
    for (auto& x : th) x.join();
}
template <typename F>
void parallel_for(size_t n, F&& fn) {
    parallel_ranges(n, layout_threads(n), [&](size_t b, size_t e, int) {
        for (size_t i = b; i < e; ++i) fn(i);
    });
}
}  // namespace

std::string build_layout(const mfs_mesh_view& v, HostLayout& L, int kTileCells) {
    if (kTileCells < 64) return "tile size must be at least 64 cells";
    if (v.nTCell <= 0 || v.nTFace <= 0 || v.nTNode <= 0 || v.nBFace < 0 || v.nBFace > v.nTFace)
        return "mesh view has invalid sizes";
    if (!v.x || !v.y || !v.z || !v.nNPF || !v.f2n || !v.f2c || !v.xfn || !v.yfn || !v.zfn || !v.area || !v.xfc ||
        !v.yfc || !v.zfc || !v.xcc || !v.ycc || !v.zcc || !v.vol || !v.rho || !v.u || !v.v || !v.w || !v.p ||
        (v.nBFace > 0 && !v.bc_type))
        return "mesh view has a null array";
    const int nC = v.nTCell, nF = v.nTFace, nB = v.nBFace, nN = v.nTNode, n = nC + nB;
    L.nC = nC;
    L.nF = nF;
    L.nB = nB;
    L.nN = nN;

    // ---- mesh-order face tables and validation
    std::vector<int> mfo(nF + 1);
    mfo[0] = 0;
    for (int i = 0; i < nF; ++i) {
        if (v.nNPF[i] < 3) return "face with fewer than 3 nodes";
        mfo[i + 1] = mfo[i] + v.nNPF[i];
    }
    for (int k = 0; k < mfo[nF]; ++k)
        if (v.f2n[k] < 0 || v.f2n[k] >= nN) return "f2n has an invalid node";
    for (int i = 0; i < nF; ++i) {
        const int a = v.f2c[2 * i], b = v.f2c[2 * i + 1];
        if (a < 0 || a >= nC) return "f2c has an invalid first cell";
        if (i < nB) {
            if (b != nC + i) return "boundary face k must have ghost cell nTCell+k (SpecifyBC)";
        } else if (b < 0 || b >= nC || b == a) {
            return "interior face has an invalid second cell";
        }
    }
    L.btype.assign(v.bc_type, v.bc_type + nB);

    // ---- mesh cell -> faces in ascending mesh face id (CalnFPC/CalC2F)
    std::vector<int> mcfo(nC + 1, 0), mcf;
    for (int i = 0; i < nF; ++i) {
        mcfo[v.f2c[2 * i] + 1]++;
        if (i >= nB) mcfo[v.f2c[2 * i + 1] + 1]++;
    }
    for (int c = 0; c < nC; ++c) mcfo[c + 1] += mcfo[c];
    mcf.resize(mcfo[nC]);
    {
        std::vector<int> fill(mcfo.begin(), mcfo.end() - 1);
        for (int i = 0; i < nF; ++i) {
            mcf[fill[v.f2c[2 * i]]++] = i;
            if (i >= nB) mcf[fill[v.f2c[2 * i + 1]]++] = i;
        }
    }
    L.max_faces = 0;
    for (int c = 0; c < nC; ++c) {
        const int k = mcfo[c + 1] - mcfo[c];
        if (k == 0) return "cell without faces";
        L.max_faces = std::max(L.max_faces, k);
    }
    if (L.max_faces > 8) return "cells with more than 8 faces are not supported";
    auto other = [&](int c, int f) { return v.f2c[2 * f] == c ? v.f2c[2 * f + 1] : v.f2c[2 * f]; };

    // ---- LU-SGS levels on mesh ids.  With LUSGSLayer[i] = i
    // (ReorderCellforLUSGS_0, grid_polyhedra.cpp:1127-1142) the forward sweep
    // of cell c reads the final DQ of every face neighbour with a smaller id
    // and the backward sweep that of every neighbour with a larger id.
    L.flevel.assign(nC, 0);
    int mf = 0;
    for (int c = 0; c < nC; ++c) {
        int l = 0;
        for (int k = mcfo[c]; k < mcfo[c + 1]; ++k) {
            const int f = mcf[k];
            if (f < nB) continue;
            const int nb = other(c, f);
            if (nb < c) l = std::max(l, L.flevel[nb] + 1);
        }
        L.flevel[c] = l;
        mf = std::max(mf, l);
    }
    L.nlev = mf + 1;

    // ---- device cell order: tile (kTileCells consecutive mesh ids, a
    // compact region in TDTree numbering), then forward level, then mesh id.
    // The cells of one level inside a tile are contiguous (sweep chunks),
    // and a cell's face neighbours are mostly in its own tile (gathers of
    // the face kernels stay in a small window of memory).
    std::vector<int> dcell(nC);
    {
        const int ntile = (nC + kTileCells - 1) / kTileCells;
        const size_t nkey = (size_t)ntile * L.nlev;
        std::vector<int> cnt(nkey + 1, 0);
        auto key = [&](int c) { return (size_t)(c / kTileCells) * L.nlev + L.flevel[c]; };
        for (int c = 0; c < nC; ++c) cnt[key(c) + 1]++;
        for (size_t k = 0; k < nkey; ++k) cnt[k + 1] += cnt[k];
        L.cell_of_dev.resize(nC);
        for (int c = 0; c < nC; ++c) {
            const int d = cnt[key(c)]++;
            dcell[c] = d;
            L.cell_of_dev[d] = c;
        }
    }
    L.dev_cell0 = dcell[0];
    auto dmap = [&](int c) { return c < nC ? dcell[c] : c; };  // ghosts keep nC+k

    // ---- device face order: first touch in the device cell traversal
    std::vector<int> dface(nF, -1);
    L.face_of_dev.resize(nF);
    {
        int next = 0;
        for (int d = 0; d < nC; ++d) {
            const int c = L.cell_of_dev[d];
            for (int k = mcfo[c]; k < mcfo[c + 1]; ++k) {
                const int f = mcf[k];
                if (dface[f] < 0) {
                    dface[f] = next;
                    L.face_of_dev[next] = f;
                    ++next;
                }
            }
        }
        if (next != nF) return "face not attached to any cell";
    }
    // per device cell: device face ids, still in ascending mesh face id,
    // and the sweep dependency masks
    L.cfo.assign(nC + 1, 0);
    L.cf.resize(mcf.size());
    L.depmask.assign(nC, 0);
    for (int d = 0; d < nC; ++d) L.cfo[d + 1] = L.cfo[d] + (mcfo[L.cell_of_dev[d] + 1] - mcfo[L.cell_of_dev[d]]);
    parallel_for(nC, [&](size_t dd) {
        {
            const int d = (int)dd;
            int pos = L.cfo[d];
            const int c = L.cell_of_dev[d];
            int s = 0;
            unsigned mask = 0;
            for (int k = mcfo[c]; k < mcfo[c + 1]; ++k, ++s) {
                const int f = mcf[k];
                L.cf[pos++] = dface[f];
                const int nb = other(c, f);
                if (nb < c) mask |= 1u << s;  // ghosts have nb >= nC > c
                if (nb > c) {
                    mask |= 1u << (8 + s);
                    if (nb < nC) mask |= 1u << (16 + s);
                }
                // gradient: boundary faces other than SYMM and INTERFACE take
                // the face-centre average of cell and ghost instead of the
                // node mean (cal_gradient.cpp:1740-1795)
                if (nb >= nC) {
                    const int t = L.btype[nb - nC];
                    if (t != BC_SYMM && t != BC_INTERFACE) mask |= 1u << (24 + s);
                }
            }
            L.depmask[d] = (int)mask;
        }
    });

    // ---- device node order: first touch over the device faces
    std::vector<int> dnode(nN, -1);
    L.node_of_dev.assign(nN, -1);
    {
        int next = 0;
        for (int d = 0; d < nF; ++d) {
            const int f = L.face_of_dev[d];
            for (int j = mfo[f]; j < mfo[f + 1]; ++j) {
                const int p = v.f2n[j];
                if (dnode[p] < 0) {
                    dnode[p] = next;
                    L.node_of_dev[next] = p;
                    ++next;
                }
            }
        }
        for (int p = 0; p < nN; ++p)  // nodes used by no face keep their relative order at the end
            if (dnode[p] < 0) {
                dnode[p] = next;
                L.node_of_dev[next] = p;
                ++next;
            }
    }

    // ---- faces in device order
    L.c1.resize(nF);
    L.c2.resize(nF);
    L.fgeo.resize(4 * (size_t)nF);
    L.dn1.resize(nF);
    L.dn2.assign(nF, 0.0);
    L.fno.resize(nF + 1);
    L.fno[0] = 0;
    for (int d = 0; d < nF; ++d) L.fno[d + 1] = L.fno[d] + v.nNPF[L.face_of_dev[d]];
    L.fn.resize(mfo[nF]);
    parallel_for(nF, [&](size_t dd) {
        const int d = (int)dd;
        const int i = L.face_of_dev[d];
        const int a = v.f2c[2 * i], b = v.f2c[2 * i + 1];
        L.c1[d] = dmap(a);
        L.c2[d] = dmap(b);
        L.fgeo[4 * (size_t)d + 0] = v.xfn[i];
        L.fgeo[4 * (size_t)d + 1] = v.yfn[i];
        L.fgeo[4 * (size_t)d + 2] = v.zfn[i];
        L.fgeo[4 * (size_t)d + 3] = v.area[i];
        // TimeStepNormal_new (compute_timestep.cpp:758-823): the same
        // expression in every iteration, so it is evaluated once here.
        L.dn1[d] = fabs((v.xfc[i] - v.xcc[a]) * v.xfn[i] + (v.yfc[i] - v.ycc[a]) * v.yfn[i] +
                        (v.zfc[i] - v.zcc[a]) * v.zfn[i]);
        if (i >= nB)
            L.dn2[d] = fabs((v.xfc[i] - v.xcc[b]) * v.xfn[i] + (v.yfc[i] - v.ycc[b]) * v.yfn[i] +
                            (v.zfc[i] - v.zcc[b]) * v.zfn[i]);
        for (int j = mfo[i], o = L.fno[d]; j < mfo[i + 1]; ++j, ++o) L.fn[o] = dnode[v.f2n[j]];
    });
    L.vol.resize(nC);
    parallel_for(nC, [&](size_t d) { L.vol[d] = v.vol[L.cell_of_dev[d]]; });

    // ---- node weights (ComputeWeight3D_Node, grid_polyhedra_auxiliary_functions.cpp:831-1049)
    // and the accumulation order of CompNodeVar3D_dist (cal_gradient.cpp:1922-2372),
    // computed on mesh ids.
    std::vector<double> wnode(nN, 0.0);
    std::vector<int> Nmark(nN, 0);
    std::vector<double> wbf(mfo[nB] + 1, 0.0);  // WeightNodeBFace2C
    auto bf_weight = [&](int i, int j, int p1) {
        const double dx = v.x[p1] - v.xfc[i], dy = v.y[p1] - v.yfc[i], dz = v.z[p1] - v.zfc[i];
        double wr = dx * dx + dy * dy + dz * dz;
        wr = sqrt(wr);
        wr = 1. / wr;
        wbf[mfo[i] + j] = wr;
        wnode[p1] += wr;
    };
    for (int i = 0; i < nB; i++) {
        if (L.btype[i] != BC_WALL) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = v.f2n[mfo[i] + j];
            bf_weight(i, j, p1);
            Nmark[p1] = BC_WALL;
        }
    }
    for (int i = 0; i < nB; i++) {
        if (L.btype[i] != BC_FAR_FIELD && L.btype[i] != BC_FAR_FIELD_MOD) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = v.f2n[mfo[i] + j];
            if (Nmark[p1] == BC_WALL) continue;
            bf_weight(i, j, p1);
            Nmark[p1] = BC_FAR_FIELD;
        }
    }
    for (int i = 0; i < nB; i++) {
        const int t = L.btype[i];
        if (t == BC_WALL || t == BC_SYMM || t == BC_FAR_FIELD || t == BC_FAR_FIELD_MOD || t == BC_INTERFACE) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = v.f2n[mfo[i] + j];
            if (Nmark[p1] == BC_WALL || Nmark[p1] == BC_FAR_FIELD) continue;
            bf_weight(i, j, p1);
            Nmark[p1] = 101;
        }
    }
    // Face loops of CompNodeVar3D_dist: WALL faces, FAR_FIELD faces (not on
    // WALL nodes), other types (not on WALL or FAR_FIELD nodes); then the
    // cell loop, whose per-node order is ascending mesh cell id because
    // each (cell,node) pair is visited once.
    auto loop_of = [&](int i) {
        const int t = L.btype[i];
        if (t == BC_WALL) return 1;
        if (t == BC_FAR_FIELD) return 2;
        if (t == BC_SYMM || t == BC_INTERFACE) return 0;
        return 3;
    };
    auto face_entry_ok = [&](int p1, int loop) {
        if (loop == 1) return true;
        if (loop == 2) return Nmark[p1] != BC_WALL;
        return Nmark[p1] != BC_WALL && Nmark[p1] != BC_FAR_FIELD;
    };
    std::vector<int> ecount(nN, 0);
    for (int loop = 1; loop <= 3; ++loop)
        for (int i = 0; i < nB; i++) {
            if (loop_of(i) != loop) continue;
            for (int j = mfo[i]; j < mfo[i + 1]; j++)
                if (face_entry_ok(v.f2n[j], loop)) ecount[v.f2n[j]]++;
        }
    // The cell loop on threads, each over a contiguous range of cells: a
    // node's entries come in ascending cell id, so thread t's entries of
    // node p follow those of threads < t (a counting sort per node).
    auto cell_nodes = [&](int c, std::vector<int>& nodes) {  // the cell's distinct nodes, first-visit order
        nodes.clear();
        for (int k = mcfo[c]; k < mcfo[c + 1]; ++k) {
            const int f = mcf[k];
            for (int j = mfo[f]; j < mfo[f + 1]; ++j) {
                const int p1 = v.f2n[j];
                if (std::find(nodes.begin(), nodes.end(), p1) == nodes.end()) nodes.push_back(p1);
            }
        }
    };
    const int TC = layout_threads(nC);
    std::vector<std::vector<int>> tcount(TC);  // per thread: interior-node entries per node, then fill offsets
    parallel_ranges(nC, TC, [&](size_t b, size_t e, int t) {
        std::vector<int>& cnt = tcount[t];
        cnt.assign(nN, 0);
        std::vector<int> nodes;
        for (int c = (int)b; c < (int)e; ++c) {
            cell_nodes(c, nodes);
            for (int p1 : nodes)
                if (Nmark[p1] == 0) cnt[p1]++;
        }
    });
    for (int t = 0; t < TC; ++t)
        for (int p1 = 0; p1 < nN; ++p1) ecount[p1] += tcount[t][p1];
    L.nlo.assign(nN + 1, 0);
    for (int d = 0; d < nN; ++d) L.nlo[d + 1] = L.nlo[d] + ecount[L.node_of_dev[d]];
    L.nla.resize(L.nlo[nN]);
    L.nlb.resize(L.nlo[nN]);
    L.nlw.resize(L.nlo[nN]);
    std::vector<int> efill(nN);
    for (int d = 0; d < nN; ++d) efill[L.node_of_dev[d]] = L.nlo[d];
    for (int loop = 1; loop <= 3; ++loop)
        for (int i = 0; i < nB; i++) {
            if (loop_of(i) != loop) continue;
            for (int j = 0; j < v.nNPF[i]; j++) {
                const int p1 = v.f2n[mfo[i] + j];
                if (!face_entry_ok(p1, loop)) continue;
                const int e = efill[p1]++;
                L.nla[e] = dmap(v.f2c[2 * i]);
                L.nlb[e] = v.f2c[2 * i + 1];  // ghost
                L.nlw[e] = wbf[mfo[i] + j];
            }
        }
    // each thread's first entry of every interior node
    parallel_for(nN, [&](size_t pp) {
        int o = efill[pp];
        for (int t = 0; t < TC; ++t) {
            const int k = tcount[t][pp];
            tcount[t][pp] = o;
            o += k;
        }
    });
    parallel_ranges(nC, TC, [&](size_t b, size_t e, int t) {
        std::vector<int>& off = tcount[t];
        std::vector<int> nodes;
        for (int c = (int)b; c < (int)e; ++c) {
            cell_nodes(c, nodes);
            for (int p1 : nodes) {
                if (Nmark[p1] != 0) continue;
                const double dx = v.x[p1] - v.xcc[c], dy = v.y[p1] - v.ycc[c], dz = v.z[p1] - v.zcc[c];
                double wr = dx * dx + dy * dy + dz * dz;
                wr = sqrt(wr);
                wr = 1. / wr;
                const int en = off[p1]++;
                L.nla[en] = dcell[c];
                L.nlb[en] = -1;
                L.nlw[en] = wr;
            }
        }
    });
    tcount.clear();
    // WeightNode of an interior node: its cell weights summed in ascending
    // cell id, the order of the reference's loop (:1007-1044)
    L.nden.resize(nN);
    parallel_for(nN, [&](size_t dd) {
        const int pp = L.node_of_dev[dd];
        if (Nmark[pp] == 0) {
            double w = 0.0;
            for (int en = L.nlo[dd]; en < L.nlo[dd + 1]; ++en) w += L.nlw[en];
            wnode[pp] = w;
        }
        L.nden[dd] = wnode[pp] + kTINY_;
    });

    // FindNodeSYMM (grid_polyhedra_auxiliary_functions.cpp:2060-2140)
    {
        std::vector<int> nodesymm(nN, 2);
        std::vector<double> xs(nN, 0.0), ys(nN, 0.0), zs(nN, 0.0), nface(nN, 0.0);
        for (int i = 0; i < nB; i++) {
            if (L.btype[i] != BC_SYMM) continue;
            for (int j = mfo[i]; j < mfo[i + 1]; j++) {
                const int p1 = v.f2n[j];
                nodesymm[p1] = 1;
                xs[p1] += v.xfn[i];
                ys[p1] += v.yfn[i];
                zs[p1] += v.zfn[i];
                nface[p1] += 1.0;
            }
        }
        L.sym_node.clear();
        L.sym_n.clear();
        for (int d = 0; d < nN; ++d) {
            const int p = L.node_of_dev[d];
            if (nodesymm[p] == 1) {
                L.sym_node.push_back(d);
                L.sym_n.push_back(xs[p] / nface[p]);
                L.sym_n.push_back(ys[p] / nface[p]);
                L.sym_n.push_back(zs[p] / nface[p]);
            }
        }
    }

    // ---- sweep chunks.  A chunk is up to `cpc` consecutive device cells of
    // one tile and one forward level, so its cells never depend on one
    // another.  Chunks are numbered in device order (tile, then level), and
    // tile k owns chunks tile_chunk[k] .. tile_chunk[k+1]-1.  A lower
    // neighbour has a smaller mesh id (same or earlier tile) and a lower
    // level, so every forward dependency lies in an earlier chunk and every
    // backward dependency in a later one.  rep_f/rep_b name, per chunk, the
    // dependency expected to be published last: the one with the highest
    // (forward) or lowest (backward) level, another tile's first.
    L.lpc = L.max_faces <= 4 ? 4 : 8;
    L.cpc = 64 / L.lpc;
    {
        std::vector<int> chunk_of(nC);
        L.chunk.clear();
        L.tile_chunk.clear();
        int cur_tile = -1;
        for (int d = 0; d < nC;) {
            const int c0 = L.cell_of_dev[d];
            const int l = L.flevel[c0], tile = c0 / kTileCells;
            while (cur_tile < tile) {
                L.tile_chunk.push_back((int)(L.chunk.size() / 2));
                ++cur_tile;
            }
            int e = d;
            while (e < nC && e - d < L.cpc && L.flevel[L.cell_of_dev[e]] == l &&
                   L.cell_of_dev[e] / kTileCells == tile)
                ++e;
            const int t = (int)(L.chunk.size() / 2);
            for (int x = d; x < e; ++x) chunk_of[x] = t;
            L.chunk.push_back(d);
            L.chunk.push_back(e - d);
            d = e;
        }
        const int nchunk = (int)(L.chunk.size() / 2);
        L.tile_chunk.push_back(nchunk);
        // flat schedule: every chunk of level l before any of level l+1,
        // chunks of one level in chunk order (a stable counting sort)
        L.run.resize(nchunk);
        {
            std::vector<int> at(L.nlev + 1, 0);
            for (int t = 0; t < nchunk; ++t) at[L.flevel[L.cell_of_dev[L.chunk[2 * t]]] + 1]++;
            for (int l = 0; l < L.nlev; ++l) at[l + 1] += at[l];
            for (int t = 0; t < nchunk; ++t) L.run[at[L.flevel[L.cell_of_dev[L.chunk[2 * t]]]]++] = t;
        }
        L.rep_f.assign(nchunk, -1);
        L.rep_b.assign(nchunk, -1);
        // rank of a dependency: level first, then "in another tile"; the
        // first of equal rank in the chunk's cell and slot order
        std::vector<char> bad(nchunk, 0);
        parallel_for(nchunk, [&](size_t tt) {
            const int t = (int)tt;
            long long bestF = -1, bestB = -1;
            for (int d = L.chunk[2 * t]; d < L.chunk[2 * t] + L.chunk[2 * t + 1]; ++d) {
                const int c = L.cell_of_dev[d];
                for (int k = mcfo[c]; k < mcfo[c + 1]; ++k) {
                    const int f = mcf[k];
                    if (f < nB) continue;
                    const int nb = other(c, f);
                    const int dn = dcell[nb];
                    const int tn = chunk_of[dn];
                    const int foreign = (nb / kTileCells != c / kTileCells) ? 1 : 0;
                    if (nb < c) {
                        if (tn >= t) bad[t] = 1;
                        const long long r = 2LL * L.flevel[nb] + foreign;
                        if (r > bestF) {
                            bestF = r;
                            L.rep_f[t] = dn;
                        }
                    } else {
                        if (tn <= t) bad[t] = 2;
                        const long long r = 2LL * (L.nlev - L.flevel[nb]) + foreign;
                        if (r > bestB) {
                            bestB = r;
                            L.rep_b[t] = dn;
                        }
                    }
                }
            }
        });
        for (int t = 0; t < nchunk; ++t)
            if (bad[t])
                return std::string(bad[t] == 1 ? "forward" : "backward") + " sweep chunks are not topological";
        // face slots per device cell (device face ids, ascending mesh face id, -1 = none)
        L.cfs.assign((size_t)nC * L.lpc, -1);
        for (int d = 0; d < nC; ++d)
            for (int k = L.cfo[d]; k < L.cfo[d + 1]; ++k) L.cfs[(size_t)d * L.lpc + (k - L.cfo[d])] = L.cf[k];
        // the sweeps' tables: chunk records by position, slot neighbours and
        // oriented geometry (the normal is negated where the cell is the
        // face's second cell, exactly as the sweep would negate it)
        // A chunk of a thin level (at most kThinLevelChunks chunks) has its
        // lanes poll their own dependencies at once instead of one lane
        // polling the latest first: few waves poll there, and the second
        // round trip is a level's latency.  Flag: bit 8 of the cell count.
        std::vector<int> lw(L.nlev, 0);
        for (int t = 0; t < nchunk; ++t) ++lw[L.flevel[L.cell_of_dev[L.chunk[2 * t]]]];
        L.crun.resize(4 * (size_t)nchunk);
        L.pos.resize(nchunk);
        for (int u = 0; u < nchunk; ++u) {
            const int t = L.run[u];
            const bool thin = lw[L.flevel[L.cell_of_dev[L.chunk[2 * t]]]] <= kThinLevelChunks;
            L.pos[t] = u;
            L.crun[4 * (size_t)u] = L.chunk[2 * t];
            L.crun[4 * (size_t)u + 1] = L.chunk[2 * t + 1] | (thin ? kThinChunkFlag : 0);
            L.crun[4 * (size_t)u + 2] = L.rep_f[t];
            L.crun[4 * (size_t)u + 3] = L.rep_b[t];
        }
        // tetrahedral meshes whose slots are all triangles take the
        // gradient's slot-node path (the device builds the slot tables from
        // cfs and the face arrays)
        L.tri_slots = false;
        if (L.max_faces == 4) {
            bool tri = true;
            for (size_t i = 0; i < L.cfs.size() && tri; ++i) {
                const int f = L.cfs[i];
                tri = f >= 0 && L.fno[f + 1] - L.fno[f] == 3;
            }
            L.tri_slots = tri;
        }
    }

    // ---- state, device order (ghosts unchanged)
    L.q.resize(5 * (size_t)n);
    const double* src[5] = {v.rho, v.u, v.v, v.w, v.p};
    parallel_for(nC, [&](size_t d) {
        for (int j = 0; j < 5; ++j) L.q[(size_t)j * n + d] = src[j][L.cell_of_dev[d]];
    });
    for (int j = 0; j < 5; ++j)
        for (int k = nC; k < n; ++k) L.q[(size_t)j * n + k] = src[j][k];
    return "";
}

}  // namespace mfs
