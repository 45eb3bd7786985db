// Device-layout construction (see mfs_prep.hpp).
#include "mfs_prep.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <sstream>

#include "mfs_mesh.hpp"

namespace mfs {

namespace {
constexpr double kTINY_ = 1.e-40;
}

std::string build_layout(const mfs_mesh_view& v, HostLayout& L) {
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

    // ---- faces
    L.c1.resize(nF);
    L.c2.resize(nF);
    L.fgeo.resize(4 * (size_t)nF);
    L.dn1.resize(nF);
    L.dn2.assign(nF, 0.0);
    L.btype.assign(v.bc_type, v.bc_type + nB);
    for (int i = 0; i < nF; ++i) {
        const int a = v.f2c[2 * i], b = v.f2c[2 * i + 1];
        if (a < 0 || a >= nC) return "f2c has an invalid first cell";
        if (i < nB) {
            if (b != nC + i) return "boundary face k must have ghost cell nTCell+k (SpecifyBC)";
        } else if (b < 0 || b >= nC || b == a) {
            return "interior face has an invalid second cell";
        }
        L.c1[i] = a;
        L.c2[i] = b;
        L.fgeo[4 * (size_t)i + 0] = v.xfn[i];
        L.fgeo[4 * (size_t)i + 1] = v.yfn[i];
        L.fgeo[4 * (size_t)i + 2] = v.zfn[i];
        L.fgeo[4 * (size_t)i + 3] = v.area[i];
        // TimeStepNormal_new (compute_timestep.cpp:758-823): the same
        // expression for every iteration, so it is evaluated once here.
        L.dn1[i] = fabs((v.xfc[i] - v.xcc[a]) * v.xfn[i] + (v.yfc[i] - v.ycc[a]) * v.yfn[i] +
                        (v.zfc[i] - v.zcc[a]) * v.zfn[i]);
        if (i >= nB)
            L.dn2[i] = fabs((v.xfc[i] - v.xcc[b]) * v.xfn[i] + (v.yfc[i] - v.ycc[b]) * v.yfn[i] +
                            (v.zfc[i] - v.zcc[b]) * v.zfn[i]);
    }
    L.fno.resize(nF + 1);
    L.fno[0] = 0;
    for (int i = 0; i < nF; ++i) {
        if (v.nNPF[i] < 3) return "face with fewer than 3 nodes";
        L.fno[i + 1] = L.fno[i] + v.nNPF[i];
    }
    L.fn.assign(v.f2n, v.f2n + L.fno[nF]);
    for (int k = 0; k < L.fno[nF]; ++k)
        if (L.fn[k] < 0 || L.fn[k] >= nN) return "f2n has an invalid node";

    // ---- cell -> faces, ascending (CalnFPC/CalC2F)
    L.cfo.assign(nC + 1, 0);
    for (int i = 0; i < nF; ++i) {
        L.cfo[L.c1[i] + 1]++;
        if (i >= nB) L.cfo[L.c2[i] + 1]++;
    }
    for (int c = 0; c < nC; ++c) L.cfo[c + 1] += L.cfo[c];
    L.cf.resize(L.cfo[nC]);
    {
        std::vector<int> fill(L.cfo.begin(), L.cfo.end() - 1);
        for (int i = 0; i < nF; ++i) {
            L.cf[fill[L.c1[i]]++] = i;
            if (i >= nB) L.cf[fill[L.c2[i]]++] = i;
        }
    }
    for (int c = 0; c < nC; ++c)
        if (L.cfo[c + 1] == L.cfo[c]) return "cell without faces";
    L.vol.assign(v.vol, v.vol + nC);

    // ---- node weights (ComputeWeight3D_Node, grid_polyhedra_auxiliary_functions.cpp:831-1049)
    // and the accumulation order of CompNodeVar3D_dist (cal_gradient.cpp:1922-2372).
    std::vector<double> wnode(nN, 0.0);
    std::vector<int>& Nmark = L.nmark;
    Nmark.assign(nN, 0);
    std::vector<double> wbf(L.fno[nF > nB ? nB : nF], 0.0);  // WeightNodeBFace2C, boundary faces only
    auto bf_weight = [&](int i, int j, int p1) {
        const double dx = v.x[p1] - v.xfc[i], dy = v.y[p1] - v.yfc[i], dz = v.z[p1] - v.zfc[i];
        double wr = dx * dx + dy * dy + dz * dz;
        wr = sqrt(wr);
        wr = 1. / wr;
        wbf[L.fno[i] + j] = wr;
        wnode[p1] += wr;
    };
    for (int i = 0; i < nB; i++) {
        if (L.btype[i] != BC_WALL) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = L.fn[L.fno[i] + j];
            bf_weight(i, j, p1);
            Nmark[p1] = BC_WALL;
        }
    }
    for (int i = 0; i < nB; i++) {
        if (L.btype[i] != BC_FAR_FIELD && L.btype[i] != BC_FAR_FIELD_MOD) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = L.fn[L.fno[i] + j];
            if (Nmark[p1] == BC_WALL) continue;
            bf_weight(i, j, p1);
            Nmark[p1] = BC_FAR_FIELD;
        }
    }
    for (int i = 0; i < nB; i++) {
        const int t = L.btype[i];
        if (t == BC_WALL || t == BC_SYMM || t == BC_FAR_FIELD || t == BC_FAR_FIELD_MOD || t == BC_INTERFACE) continue;
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = L.fn[L.fno[i] + j];
            if (Nmark[p1] == BC_WALL || Nmark[p1] == BC_FAR_FIELD) continue;
            bf_weight(i, j, p1);
            Nmark[p1] = 101;
        }
    }
    // cell -> node list of each cell: the nodes of its faces, unique.  Each
    // (cell,node) pair is visited once by the reference's cell loop, whose
    // per-node summation order is therefore ascending cell id.
    std::vector<std::vector<int>> c2n(nC);
    for (int c = 0; c < nC; ++c) {
        for (int k = L.cfo[c]; k < L.cfo[c + 1]; ++k) {
            const int f = L.cf[k];
            for (int j = L.fno[f]; j < L.fno[f + 1]; ++j) c2n[c].push_back(L.fn[j]);
        }
        std::sort(c2n[c].begin(), c2n[c].end());
        c2n[c].erase(std::unique(c2n[c].begin(), c2n[c].end()), c2n[c].end());
    }
    // entries grouped by node, in reference order
    struct Ent {
        int a, b;
        double w;
    };
    std::vector<std::vector<Ent>> ent(nN);
    auto add_face_entries = [&](int i, bool (*skip)(int)) {
        for (int j = 0; j < v.nNPF[i]; j++) {
            const int p1 = L.fn[L.fno[i] + j];
            if (skip(Nmark[p1])) continue;
            ent[p1].push_back({L.c1[i], L.c2[i], wbf[L.fno[i] + j]});
        }
    };
    for (int i = 0; i < nB; i++)
        if (L.btype[i] == BC_WALL) add_face_entries(i, [](int) { return false; });
    for (int i = 0; i < nB; i++)
        if (L.btype[i] == BC_FAR_FIELD) add_face_entries(i, [](int m) { return m == BC_WALL; });
    for (int i = 0; i < nB; i++) {
        const int t = L.btype[i];
        if (t == BC_WALL || t == BC_SYMM || t == BC_FAR_FIELD || t == BC_INTERFACE) continue;
        add_face_entries(i, [](int m) { return m == BC_WALL || m == BC_FAR_FIELD; });
    }
    for (int c = 0; c < nC; ++c)
        for (int p1 : c2n[c]) {
            if (Nmark[p1] != 0) continue;
            const double dx = v.x[p1] - v.xcc[c], dy = v.y[p1] - v.ycc[c], dz = v.z[p1] - v.zcc[c];
            double wr = dx * dx + dy * dy + dz * dz;
            wr = sqrt(wr);
            wr = 1. / wr;
            wnode[p1] += wr;
            ent[p1].push_back({c, -1, wr});
        }
    std::vector<std::vector<int>>().swap(c2n);
    L.nlo.assign(nN + 1, 0);
    for (int p = 0; p < nN; ++p) L.nlo[p + 1] = L.nlo[p] + (int)ent[p].size();
    L.nla.resize(L.nlo[nN]);
    L.nlb.resize(L.nlo[nN]);
    L.nlw.resize(L.nlo[nN]);
    for (int p = 0; p < nN; ++p) {
        int k = L.nlo[p];
        for (const Ent& e : ent[p]) {
            L.nla[k] = e.a;
            L.nlb[k] = e.b;
            L.nlw[k] = e.w;
            ++k;
        }
    }
    std::vector<std::vector<Ent>>().swap(ent);
    L.nden.resize(nN);
    for (int p = 0; p < nN; ++p) L.nden[p] = wnode[p] + kTINY_;

    // FindNodeSYMM (grid_polyhedra_auxiliary_functions.cpp:2060-2140)
    {
        std::vector<int> nodesymm(nN, 2);
        std::vector<double> xs(nN, 0.0), ys(nN, 0.0), zs(nN, 0.0), nface(nN, 0.0);
        for (int i = 0; i < nB; i++) {
            if (L.btype[i] != BC_SYMM) continue;
            for (int j = L.fno[i]; j < L.fno[i + 1]; j++) {
                const int p1 = L.fn[j];
                nodesymm[p1] = 1;
                xs[p1] += v.xfn[i];
                ys[p1] += v.yfn[i];
                zs[p1] += v.zfn[i];
                nface[p1] += 1.0;
            }
        }
        L.sym_node.clear();
        L.sym_n.clear();
        for (int p = 0; p < nN; ++p)
            if (nodesymm[p] == 1) {
                L.sym_node.push_back(p);
                L.sym_n.push_back(xs[p] / nface[p]);
                L.sym_n.push_back(ys[p] / nface[p]);
                L.sym_n.push_back(zs[p] / nface[p]);
            }
    }

    // ---- LU-SGS level schedule.  With LUSGSLayer[i] = i
    // (ReorderCellforLUSGS_0, grid_polyhedra.cpp:1127-1142) the forward sweep
    // of cell c reads the final DQ of every face neighbour with a smaller id
    // and the backward sweep that of every neighbour with a larger id.
    L.flevel.assign(nC, 0);
    L.blevel.assign(nC, 0);
    int mf = 0, mb = 0;
    for (int c = 0; c < nC; ++c) {
        int l = 0;
        for (int k = L.cfo[c]; k < L.cfo[c + 1]; ++k) {
            const int f = L.cf[k];
            if (f < nB) continue;
            const int nb = (L.c1[f] == c) ? L.c2[f] : L.c1[f];
            if (nb < c) l = std::max(l, L.flevel[nb] + 1);
        }
        L.flevel[c] = l;
        mf = std::max(mf, l);
    }
    for (int c = nC - 1; c >= 0; --c) {
        int l = 0;
        for (int k = L.cfo[c]; k < L.cfo[c + 1]; ++k) {
            const int f = L.cf[k];
            if (f < nB) continue;
            const int nb = (L.c1[f] == c) ? L.c2[f] : L.c1[f];
            if (nb > c) l = std::max(l, L.blevel[nb] + 1);
        }
        L.blevel[c] = l;
        mb = std::max(mb, l);
    }
    L.nlev_f = mf + 1;
    L.nlev_b = mb + 1;
    auto bucket = [nC](const std::vector<int>& lev, int nlev, bool descending_id, std::vector<int>& out) {
        std::vector<int> cnt(nlev + 1, 0);
        for (int c = 0; c < nC; ++c) cnt[lev[c] + 1]++;
        for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
        out.resize(nC);
        if (!descending_id) {
            for (int c = 0; c < nC; ++c) out[cnt[lev[c]]++] = c;
        } else {
            for (int c = nC - 1; c >= 0; --c) out[cnt[lev[c]]++] = c;
        }
    };
    bucket(L.flevel, L.nlev_f, false, L.sched_f);
    bucket(L.blevel, L.nlev_b, true, L.sched_b);

    // ---- state
    L.q.resize(5 * (size_t)n);
    const double* src[5] = {v.rho, v.u, v.v, v.w, v.p};
    for (int j = 0; j < 5; ++j) std::copy(src[j], src[j] + n, L.q.begin() + (size_t)j * n);
    return "";
}

}  // namespace mfs
