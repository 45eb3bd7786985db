// TDTree decomposition + renumbering (see mfs_tdtree.hpp for the reference map).
//
// Rows of the cell tables (c2c, c2f, c2n) are kept under original cell ids and
// a position<->id map is permuted instead of the rows themselves; the values
// stored in the rows are original ids, as in the reference, where
// TDTree_permute moves rows but never renames their entries
// (TDTree_permutations.cpp:7-33).
#include "mfs_tdtree.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>

namespace mfs {

namespace {

constexpr int ISO = 1000000000;  // the reference's 1e9 separator tag

struct Csr {
    std::vector<int> off, val;
    int row_size(int r) const { return off[r + 1] - off[r]; }
    const int* row(int r) const { return val.data() + off[r]; }
};

class Builder {
   public:
    Builder(const RawMesh& m, int nbParts, int partSize) : m_(m), nbParts_(nbParts), partSize_(partSize) {
        t_.nbParts = nbParts;
        t_.partSize = partSize;
        const int nC = m.nC, nF = m.nF, nN = m.nN;
        f2nOff_ = m.f2n_offsets();
        // c2n / c2f / c2c as unique ascending lists (CreateTree, grid_polyhedra.cpp:376-525)
        std::vector<std::vector<int>> c2n(nC), c2f(nC), c2c(nC);
        for (int i = 0; i < nF; ++i) {
            const int c1 = m.f2c[2 * i], c2 = m.f2c[2 * i + 1];
            const bool interior = (i >= m.nB);
            for (int k = f2nOff_[i]; k < f2nOff_[i + 1]; ++k) {
                c2n[c1].push_back(m.f2n[k]);
                if (interior) c2n[c2].push_back(m.f2n[k]);
            }
            c2f[c1].push_back(i);
            if (interior) {
                c2f[c2].push_back(i);
                c2c[c1].push_back(c1);
                c2c[c2].push_back(c2);
                c2c[c1].push_back(c2);
                c2c[c2].push_back(c1);
            }
        }
        auto pack = [nC](std::vector<std::vector<int>>& rows, Csr& out) {
            out.off.assign(nC + 1, 0);
            for (int c = 0; c < nC; ++c) {
                std::sort(rows[c].begin(), rows[c].end());
                rows[c].erase(std::unique(rows[c].begin(), rows[c].end()), rows[c].end());
                out.off[c + 1] = out.off[c] + (int)rows[c].size();
            }
            out.val.resize(out.off[nC]);
            for (int c = 0; c < nC; ++c) std::copy(rows[c].begin(), rows[c].end(), out.val.begin() + out.off[c]);
            std::vector<std::vector<int>>().swap(rows);
        };
        pack(c2n, c2n_);
        pack(c2f, c2f_);
        pack(c2c, c2c_);

        t_.cellRev.resize(nC);
        t_.cellPerm.resize(nC);
        t_.faceRev.resize(nF);
        t_.facePerm.resize(nF);
        t_.nodeRev.resize(nN);
        t_.nodePerm.resize(nN);
        std::iota(t_.cellRev.begin(), t_.cellRev.end(), 0);
        std::iota(t_.cellPerm.begin(), t_.cellPerm.end(), 0);
        std::iota(t_.faceRev.begin(), t_.faceRev.end(), 0);
        std::iota(t_.facePerm.begin(), t_.facePerm.end(), 0);
        std::iota(t_.nodeRev.begin(), t_.nodeRev.end(), 0);
        std::iota(t_.nodePerm.begin(), t_.nodePerm.end(), 0);

        nodeStamp_.assign(nN, -1);
        nodeLocal_.assign(nN, -1);
        markStamp_.assign(nC, -1);
        markPart_.assign(nC, 0);
    }

    TDTree run() {
        const int nC = m_.nC, nF = m_.nF, nN = m_.nN, nB = m_.nB;
        const int nPFace = nB;  // serial meshes: no interface faces
        // TDTree_creation (TDTree_creation.cpp:407-519)
        int root = new_node();
        init_node(root, 0, nC - 1, 0, nF - 1, 0, nN - 1, false);
        t_.nodes[root].isLeaf = false;
        t_.nodes[root].nbParts = 1;
        t_.nodes[root].son.assign(1, -1);
        if (nPFace > 0) {
            int s = new_node();
            t_.nodes[root].son[0] = s;
            create_boundary(s, 0, nPFace - 1, 0, -1, false);
        }
        int iso = new_node();
        t_.nodes[root].iso = iso;
        init_node(iso, 0, nC - 1, nPFace, nF - 1, 0, nN - 1, false);
        t_.nodes[iso].isLeaf = false;
        t_.nodes[iso].nbParts = 1;
        t_.nodes[iso].son.assign(1, -1);  // no interface faces in a serial mesh
        int cells = new_node();
        t_.nodes[iso].iso = cells;
        create_normal(cells, 0, nC - 1, nB, nF - 1, 0, nN - 1, false);
        for (int i = 0; i < nC; ++i) t_.cellPerm[t_.cellRev[i]] = i;
        for (int i = 0; i < nF; ++i) t_.facePerm[t_.faceRev[i]] = i;
        for (int i = 0; i < nN; ++i) t_.nodePerm[t_.nodeRev[i]] = i;
        return std::move(t_);
    }

   private:
    const RawMesh& m_;
    int nbParts_, partSize_;
    TDTree t_;
    std::vector<int> f2nOff_;
    Csr c2n_, c2f_, c2c_;
    std::vector<int> nodeStamp_, nodeLocal_;
    std::vector<int> markStamp_, markPart_;
    int stamp_ = 0;

    int new_node() {
        t_.nodes.emplace_back();
        return (int)t_.nodes.size() - 1;
    }
    void init_node(int n, int fc, int lc, int ff, int lf, int fn, int ln, bool isIso) {
        TDTreeNode& d = t_.nodes[n];
        d.firstCell = fc;
        d.lastCell = lc;
        d.firstFace = ff;
        d.lastFace = lf;
        d.firstNode = fn;
        d.lastNode = ln;
        d.isIso = isIso;
        d.isLeaf = true;
        d.nbParts = 0;
    }

    // Recursive coordinate bisection of the nodes into nparts parts
    // (stands in for METIS_PartGraphRecursive on the nodal graph).
    void rcb(std::vector<int>& ids, int lo, int hi, int part0, int nparts, std::vector<int>& partOfLocal) {
        if (nparts <= 1 || hi - lo <= 1) {
            for (int k = lo; k < hi; ++k) partOfLocal[nodeLocal_[ids[k]]] = part0;
            return;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int k = lo; k < hi; ++k) {
            const int n = ids[k];
            const double c[3] = {m_.x[n], m_.y[n], m_.z[n]};
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], c[a]);
                mx[a] = std::max(mx[a], c[a]);
            }
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
        const std::vector<double>& coord = axis == 0 ? m_.x : (axis == 1 ? m_.y : m_.z);
        const int nleft = nparts / 2;
        const int mid = lo + (int)(((long long)(hi - lo) * nleft) / nparts);
        std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](int a, int b) {
            if (coord[a] != coord[b]) return coord[a] < coord[b];
            return a < b;
        });
        rcb(ids, lo, mid, part0, nleft, partOfLocal);
        rcb(ids, mid, hi, part0 + nleft, nparts - nleft, partOfLocal);
    }

    // Local node numbering by first appearance (create_local_i2n,
    // TDTree_partitioning.cpp:43-58) and partition of those nodes.
    // Returns the number of local nodes; nodeLocal_ holds the local ids.
    int partition_items(const std::vector<const int*>& rows, const std::vector<int>& lens, std::vector<int>& partOfLocal) {
        ++stamp_;
        std::vector<int> ids;
        for (size_t r = 0; r < rows.size(); ++r)
            for (int k = 0; k < lens[r]; ++k) {
                const int n = rows[r][k];
                if (nodeStamp_[n] != stamp_) {
                    nodeStamp_[n] = stamp_;
                    nodeLocal_[n] = (int)ids.size();
                    ids.push_back(n);
                }
            }
        partOfLocal.assign(ids.size(), 0);
        rcb(ids, 0, (int)ids.size(), 0, nbParts_, partOfLocal);
        return (int)ids.size();
    }

    // TDTree_create_itemPart (TDTree_creation.cpp:58-85): an item whose nodes
    // all share a part joins it, any other item is a separator item.
    int item_part(const int* row, int len, const std::vector<int>& partOfLocal) {
        int a = -1, b = -1;
        for (int k = 0; k < len; ++k) {
            const int p = partOfLocal[nodeLocal_[row[k]]];
            if (a == -1)
                a = p;
            else if (a != p)
                b = p;
        }
        return b == -1 ? a : ISO;
    }

    // Stable sort of the items at positions [first, first+n) by part,
    // separator last (TDTree_create_permutation + merge_permutations).
    static void stable_by_part(std::vector<int>& rev, int first, const std::vector<int>& part) {
        const int n = (int)part.size();
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return part[a] < part[b]; });
        std::vector<int> tmp(n);
        for (int k = 0; k < n; ++k) tmp[k] = rev[first + order[k]];
        std::copy(tmp.begin(), tmp.end(), rev.begin() + first);
    }

    // TDTree_create_itemPart_accordingly (TDTree_creation.cpp:88-128): items
    // (faces or nodes) in [fi, li] inherit the part of their cells; items of
    // separator cells are separator items.
    void items_accordingly(const Csr& c2i, const std::vector<int>& itemRev, int fc, const std::vector<int>& cellPart,
                           int fi, int li, std::vector<int>& itemPart, std::vector<int>& cnt, int& nIso) {
        const int nItem = li - fi + 1;
        itemPart.assign(std::max(nItem, 0), ISO);
        if (nItem <= 0) return;
        // position lookup restricted to the range
        ++stamp_;
        std::vector<int>& pos = scratchPos_;
        if ((int)pos.size() < (int)itemRev.size()) pos.assign(itemRev.size(), -1), posStamp_.assign(itemRev.size(), -1);
        for (int k = fi; k <= li; ++k) {
            pos[itemRev[k]] = k - fi;
            posStamp_[itemRev[k]] = stamp_;
        }
        const int nCell = (int)cellPart.size();
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] == ISO) continue;
            const int c = t_.cellRev[fc + i];
            for (int k = c2i.off[c]; k < c2i.off[c + 1]; ++k) {
                const int it = c2i.val[k];
                if (posStamp_[it] != stamp_) continue;
                itemPart[pos[it]] = cellPart[i];
            }
        }
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] != ISO) continue;
            const int c = t_.cellRev[fc + i];
            for (int k = c2i.off[c]; k < c2i.off[c + 1]; ++k) {
                const int it = c2i.val[k];
                if (posStamp_[it] != stamp_) continue;
                itemPart[pos[it]] = ISO;
            }
        }
        for (int k = 0; k < nItem; ++k) {
            if (itemPart[k] == ISO)
                ++nIso;
            else
                ++cnt[itemPart[k]];
        }
    }
    std::vector<int> scratchPos_, posStamp_;

    // TDTree::TDTree_create_boundary (TDTree_creation.cpp:131-262)
    void create_boundary(int tree, int ff, int lf, int fn, int ln, bool isIso) {
        const int nFace = lf - ff + 1;
        if (nFace < partSize_) {
            init_node(tree, 0, -1, ff, lf, fn, ln, isIso);
            return;
        }
        std::vector<const int*> rows(nFace);
        std::vector<int> lens(nFace);
        for (int i = 0; i < nFace; ++i) {
            const int f = t_.faceRev[ff + i];
            rows[i] = m_.f2n.data() + f2nOff_[f];
            lens[i] = f2nOff_[f + 1] - f2nOff_[f];
        }
        std::vector<int> partOfLocal;
        partition_items(rows, lens, partOfLocal);
        const int parts = nbParts_;
        std::vector<int> facePart(nFace), nbPartFace(parts, 0);
        int nbIsoFace = 0;
        for (int i = 0; i < nFace; ++i) {
            facePart[i] = item_part(rows[i], lens[i], partOfLocal);
            if (facePart[i] == ISO)
                ++nbIsoFace;
            else
                ++nbPartFace[facePart[i]];
        }
        // faces of one owner cell stay in one part (TDTree_creation.cpp:160-175)
        ++stamp_;
        for (int i = 0; i < nFace; ++i) {
            if (facePart[i] == ISO) continue;
            const int c1 = m_.f2c[2 * t_.faceRev[ff + i]];
            if (markStamp_[c1] != stamp_ || markPart_[c1] == facePart[i]) {
                markStamp_[c1] = stamp_;
                markPart_[c1] = facePart[i];
            } else {
                --nbPartFace[facePart[i]];
                facePart[i] = ISO;
                ++nbIsoFace;
            }
        }
        std::vector<int> nbPartNode(parts, 0);
        int nbIsoNode = 0;
        if (ln - fn + 1 > 0) throw std::logic_error("boundary face tree with a node range");
        stable_by_part(t_.faceRev, ff, facePart);

        const int nbIsoItem = std::max(nbIsoFace, nbIsoNode);
        std::vector<int> nbPartItem(parts);
        for (int i = 0; i < parts; ++i) nbPartItem[i] = std::max(nbPartFace[i], nbPartNode[i]);
        init_node(tree, 0, -1, ff, lf, fn, ln, isIso);
        t_.nodes[tree].nbParts = parts;
        t_.nodes[tree].son.assign(parts, -1);
        for (int i = 0; i < parts; ++i)
            if (nbPartItem[i] > 0) {
                int s = new_node();
                t_.nodes[tree].son[i] = s;
            }
        if (nbIsoItem > 0) {
            int s = new_node();
            t_.nodes[tree].iso = s;
        }
        bool leaf = true;
        int st2 = ff;
        for (int i = 0; i < parts; ++i) {
            const int ed2 = st2 + nbPartFace[i] - 1;
            if (nbPartItem[i] > 0) {
                leaf = false;
                create_boundary(t_.nodes[tree].son[i], st2, ed2, 0, -1, isIso);
            }
            st2 = ed2 + 1;
        }
        t_.nodes[tree].isLeaf = leaf;
        if (nbIsoItem > 0 && !leaf) create_boundary(t_.nodes[tree].iso, lf - nbIsoFace + 1, lf, 0, -1, true);
    }

    // TDTree::TDTree_create_normal (TDTree_creation.cpp:264-405)
    void create_normal(int tree, int fc, int lc, int ff, int lf, int fn, int ln, bool isIso) {
        const int nCell = lc - fc + 1;
        if (nCell < partSize_) {
            init_node(tree, fc, lc, ff, lf, fn, ln, isIso);
            return;
        }
        std::vector<const int*> rows(nCell);
        std::vector<int> lens(nCell);
        for (int i = 0; i < nCell; ++i) {
            const int c = t_.cellRev[fc + i];
            rows[i] = c2n_.row(c);
            lens[i] = c2n_.row_size(c);
        }
        std::vector<int> partOfLocal;
        partition_items(rows, lens, partOfLocal);
        const int parts = nbParts_;
        std::vector<int> cellPart(nCell), nbPartCell(parts, 0);
        int nbIsoCell = 0;
        for (int i = 0; i < nCell; ++i) {
            cellPart[i] = item_part(rows[i], lens[i], partOfLocal);
            if (cellPart[i] == ISO)
                ++nbIsoCell;
            else
                ++nbPartCell[cellPart[i]];
        }
        // a cell whose neighbourhood was claimed by another part becomes a
        // separator cell (TDTree_creation.cpp:293-313)
        ++stamp_;
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] == ISO) continue;
            const int c = t_.cellRev[fc + i];
            const int* nb = c2c_.row(c);
            const int nn = c2c_.row_size(c);
            for (int j = 0; j < nn; ++j) {
                const int x = nb[j];
                if (markStamp_[x] == stamp_ && markPart_[x] != cellPart[i]) {
                    --nbPartCell[cellPart[i]];
                    ++nbIsoCell;
                    cellPart[i] = ISO;
                    break;
                }
            }
            if (cellPart[i] == ISO) continue;
            for (int j = 0; j < nn; ++j) {
                markStamp_[nb[j]] = stamp_;
                markPart_[nb[j]] = cellPart[i];
            }
        }
        std::vector<int> facePart, nodePart, nbPartFace(parts, 0), nbPartNode(parts, 0);
        int nbIsoFace = 0, nbIsoNode = 0;
        items_accordingly(c2f_, t_.faceRev, fc, cellPart, ff, lf, facePart, nbPartFace, nbIsoFace);
        if (lf - ff + 1 > 0) stable_by_part(t_.faceRev, ff, facePart);
        items_accordingly(c2n_, t_.nodeRev, fc, cellPart, fn, ln, nodePart, nbPartNode, nbIsoNode);
        if (ln - fn + 1 > 0) stable_by_part(t_.nodeRev, fn, nodePart);
        stable_by_part(t_.cellRev, fc, cellPart);

        const int nbIsoItem = std::max(nbIsoCell, std::max(nbIsoFace, nbIsoNode));
        std::vector<int> nbPartItem(parts);
        for (int i = 0; i < parts; ++i) nbPartItem[i] = std::max(nbPartCell[i], std::max(nbPartFace[i], nbPartNode[i]));
        init_node(tree, fc, lc, ff, lf, fn, ln, isIso);
        t_.nodes[tree].nbParts = parts;
        t_.nodes[tree].son.assign(parts, -1);
        for (int i = 0; i < parts; ++i)
            if (nbPartItem[i] > 0) {
                int s = new_node();
                t_.nodes[tree].son[i] = s;
            }
        if (nbIsoItem > 0) {
            int s = new_node();
            t_.nodes[tree].iso = s;
        }
        bool leaf = true;
        int st1 = fc, st2 = ff, st3 = fn;
        for (int i = 0; i < parts; ++i) {
            const int ed1 = st1 + nbPartCell[i] - 1;
            const int ed2 = st2 + nbPartFace[i] - 1;
            const int ed3 = st3 + nbPartNode[i] - 1;
            if (nbPartItem[i] > 0) {
                leaf = false;
                create_normal(t_.nodes[tree].son[i], st1, ed1, st2, ed2, st3, ed3, isIso);
            }
            st1 = ed1 + 1;
            st2 = ed2 + 1;
            st3 = ed3 + 1;
        }
        t_.nodes[tree].isLeaf = leaf;
        if (nbIsoItem > 0 && !leaf)
            create_normal(t_.nodes[tree].iso, lc - nbIsoCell + 1, lc, lf - nbIsoFace + 1, lf, ln - nbIsoNode + 1, ln, true);
        else
            t_.nodes[tree].nbParts = 0;
    }
};

}  // namespace

void TDTree::leaves_forward(int node, std::vector<int>& out) const {
    const TDTreeNode& d = nodes[node];
    if (d.isLeaf) {
        out.push_back(node);
        return;
    }
    for (int s : d.son)
        if (s >= 0) leaves_forward(s, out);
    if (d.iso >= 0) leaves_forward(d.iso, out);
}

TDTree build_tdtree(const RawMesh& m, int nbParts, int partSize) {
    Builder b(m, nbParts, partSize);
    return b.run();
}

RawMesh renumber(const RawMesh& m, const TDTree& t) {
    RawMesh r;
    r.nN = m.nN;
    r.nF = m.nF;
    r.nC = m.nC;
    r.nB = m.nB;
    r.bc_patch = m.bc_patch;
    r.bc_type = m.bc_type;
    // TDTree_NodeReordering (grid_polyhedra.cpp:278-353)
    r.x.resize(m.nN);
    r.y.resize(m.nN);
    r.z.resize(m.nN);
    for (int i = 0; i < m.nN; ++i) {
        r.x[i] = m.x[t.nodeRev[i]];
        r.y[i] = m.y[t.nodeRev[i]];
        r.z[i] = m.z[t.nodeRev[i]];
    }
    // TDTree_CellReordering + TDTree_FaceReordering (grid_polyhedra.cpp:88-276)
    std::vector<int> off = m.f2n_offsets();
    r.nNPF.resize(m.nF);
    r.f2c.resize(2 * (size_t)m.nF);
    r.f2n.clear();
    r.f2n.reserve(m.f2n.size());
    for (int i = 0; i < m.nF; ++i) {
        const int o = t.faceRev[i];
        r.nNPF[i] = m.nNPF[o];
        for (int k = off[o]; k < off[o + 1]; ++k) r.f2n.push_back(t.nodePerm[m.f2n[k]]);
        const int c1 = m.f2c[2 * o], c2 = m.f2c[2 * o + 1];
        r.f2c[2 * i] = (c1 >= 0 && c1 < m.nC) ? t.cellPerm[c1] : c1;
        r.f2c[2 * i + 1] = (c2 >= 0 && c2 < m.nC) ? t.cellPerm[c2] : c2;
    }
    return r;
}

}  // namespace mfs
