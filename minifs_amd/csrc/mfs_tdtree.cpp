// TDTree decomposition + renumbering (see mfs_tdtree.hpp for the reference map).
//
// Rows of the cell tables (c2c, c2f, c2n) are kept under original cell ids and
// a position<->id map is permuted instead of the rows themselves; the values
// stored in the rows are original ids, as in the reference, where
// TDTree_permute moves rows but never renames their entries
// (TDTree_permutations.cpp:7-33).
#include "mfs_tdtree.hpp"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mfs {

namespace {

constexpr int ISO = 1000000000;  // the reference's 1e9 separator tag

struct Csr {
    std::vector<int> off, val;
    int row_size(int r) const { return off[r + 1] - off[r]; }
    const int* row(int r) const { return val.data() + off[r]; }
};

// Subtrees of at least this many cells build their parts on other threads.
constexpr int kParallelCells = 32768;

// Per-thread marks of the recursion.  Each call takes a fresh stamp, so an
// entry counts only when its stamp is the current call's and nothing needs
// clearing between calls.
struct Scratch {
    std::vector<int> nodeStamp, nodeLocal;  // by node: local numbering of a call's nodes
    std::vector<int> markStamp, markPart;   // by cell: parts claimed by neighbours
    std::vector<int> pos, posStamp;         // by face or node: position inside a call's range
    int stamp = 0;
    Scratch(int nN, int nC) : nodeStamp(nN, -1), nodeLocal(nN, -1), markStamp(nC, -1), markPart(nC, 0) {}
};

// What the subtrees share: the mesh, the cell tables and the permutations
// (each subtree writes only its own cell, face and node ranges).
struct Shared {
    const RawMesh& m;
    int nbParts, partSize;
    TDTree t;
    std::vector<int> f2nOff;
    Csr c2n, c2f, c2c;
    std::atomic<int> spare{0};  // threads that may still be started
    std::mutex pool_mu;
    std::vector<std::unique_ptr<Scratch>> pool;  // scratch of finished workers, reused

    std::unique_ptr<Scratch> get_scratch() {
        {
            std::lock_guard<std::mutex> lk(pool_mu);
            if (!pool.empty()) {
                std::unique_ptr<Scratch> sc = std::move(pool.back());
                pool.pop_back();
                return sc;
            }
        }
        return std::make_unique<Scratch>(m.nN, m.nC);
    }
    void put_scratch(std::unique_ptr<Scratch> sc) {
        std::lock_guard<std::mutex> lk(pool_mu);
        pool.push_back(std::move(sc));
    }

    Shared(const RawMesh& m_, int nbParts_, int partSize_) : m(m_), nbParts(nbParts_), partSize(partSize_) {
        t.nbParts = nbParts_;
        t.partSize = partSize_;
        const int nC = m.nC, nF = m.nF, nN = m.nN;
        f2nOff = m.f2n_offsets();
        // c2n / c2f / c2c as unique ascending lists (CreateTree,
        // grid_polyhedra.cpp:376-525): rows counted, filled in face order,
        // then each row sorted and deduplicated in place
        std::vector<int> nn(nC, 0), nf(nC, 0), ncc(nC, 0);
        for (int i = 0; i < nF; ++i) {
            const int c1 = m.f2c[2 * i], c2 = m.f2c[2 * i + 1];
            const int k = f2nOff[i + 1] - f2nOff[i];
            nn[c1] += k;
            ++nf[c1];
            if (i >= m.nB) {
                nn[c2] += k;
                ++nf[c2];
                ncc[c1] += 2;
                ncc[c2] += 2;
            }
        }
        auto open_rows = [nC](const std::vector<int>& cnt, Csr& out) {
            out.off.assign(nC + 1, 0);
            for (int c = 0; c < nC; ++c) out.off[c + 1] = out.off[c] + cnt[c];
            out.val.resize(out.off[nC]);
        };
        open_rows(nn, c2n);
        open_rows(nf, c2f);
        open_rows(ncc, c2c);
        for (int c = 0; c < nC; ++c) nn[c] = c2n.off[c], nf[c] = c2f.off[c], ncc[c] = c2c.off[c];
        for (int i = 0; i < nF; ++i) {
            const int c1 = m.f2c[2 * i], c2 = m.f2c[2 * i + 1];
            const bool interior = (i >= m.nB);
            for (int k = f2nOff[i]; k < f2nOff[i + 1]; ++k) {
                c2n.val[nn[c1]++] = m.f2n[k];
                if (interior) c2n.val[nn[c2]++] = m.f2n[k];
            }
            c2f.val[nf[c1]++] = i;
            if (interior) {
                c2f.val[nf[c2]++] = i;
                c2c.val[ncc[c1]++] = c1;
                c2c.val[ncc[c1]++] = c2;
                c2c.val[ncc[c2]++] = c2;
                c2c.val[ncc[c2]++] = c1;
            }
        }
        auto close_rows = [nC](Csr& t) {
            int w = 0;
            for (int c = 0; c < nC; ++c) {
                int* const b = t.val.data() + t.off[c];
                int* const e = t.val.data() + t.off[c + 1];
                std::sort(b, e);
                int* const u = std::unique(b, e);
                t.off[c] = w;
                for (int* q = b; q < u; ++q) t.val[w++] = *q;
            }
            t.off[nC] = w;
            t.val.resize(w);
            t.val.shrink_to_fit();
        };
        close_rows(c2n);
        close_rows(c2f);
        close_rows(c2c);

        t.cellRev.resize(nC);
        t.cellPerm.resize(nC);
        t.faceRev.resize(nF);
        t.facePerm.resize(nF);
        t.nodeRev.resize(nN);
        t.nodePerm.resize(nN);
        std::iota(t.cellRev.begin(), t.cellRev.end(), 0);
        std::iota(t.cellPerm.begin(), t.cellPerm.end(), 0);
        std::iota(t.faceRev.begin(), t.faceRev.end(), 0);
        std::iota(t.facePerm.begin(), t.facePerm.end(), 0);
        std::iota(t.nodeRev.begin(), t.nodeRev.end(), 0);
        std::iota(t.nodePerm.begin(), t.nodePerm.end(), 0);

            const unsigned hw = std::thread::hardware_concurrency();
        spare = (int)std::min(8u, std::max(1u, hw)) - 1;
    }
    bool take_thread() {
        int v = spare.load();
        while (v > 0)
            if (spare.compare_exchange_weak(v, v - 1)) return true;
        return false;
    }
};

class Worker {
   public:
    Worker(Shared& s, std::vector<TDTreeNode>* nodes) : S(s), N(nodes), own_(s.get_scratch()), sc_(*own_) {}
    ~Worker() { S.put_scratch(std::move(own_)); }

    void run_root() {
        const int nC = S.m.nC, nF = S.m.nF, nN = S.m.nN, nB = S.m.nB;
        const int nPFace = nB;  // serial meshes: no interface faces
        // TDTree_creation (TDTree_creation.cpp:407-519)
        int root = new_node();
        init_node(root, 0, nC - 1, 0, nF - 1, 0, nN - 1, false);
        (*N)[root].isLeaf = false;
        (*N)[root].nbParts = 1;
        (*N)[root].son.assign(1, -1);
        if (nPFace > 0) {
            int s = new_node();
            (*N)[root].son[0] = s;
            create_boundary(s, 0, nPFace - 1, 0, -1, false);
        }
        int iso = new_node();
        (*N)[root].iso = iso;
        init_node(iso, 0, nC - 1, nPFace, nF - 1, 0, nN - 1, false);
        (*N)[iso].isLeaf = false;
        (*N)[iso].nbParts = 1;
        (*N)[iso].son.assign(1, -1);  // no interface faces in a serial mesh
        int cells = new_node();
        (*N)[iso].iso = cells;
        create_normal(cells, 0, nC - 1, nB, nF - 1, 0, nN - 1, false);
    }

   private:
    Shared& S;
    std::vector<TDTreeNode>* N;  // the node list this worker appends to
    std::unique_ptr<Scratch> own_;
    Scratch& sc_;

    int new_node() {
        (*N).emplace_back();
        return (int)(*N).size() - 1;
    }
    void init_node(int n, int fc, int lc, int ff, int lf, int fn, int ln, bool isIso) {
        TDTreeNode& d = (*N)[n];
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
            for (int k = lo; k < hi; ++k) partOfLocal[sc_.nodeLocal[ids[k]]] = part0;
            return;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int k = lo; k < hi; ++k) {
            const int n = ids[k];
            const double c[3] = {S.m.x[n], S.m.y[n], S.m.z[n]};
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], c[a]);
                mx[a] = std::max(mx[a], c[a]);
            }
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
        const std::vector<double>& coord = axis == 0 ? S.m.x : (axis == 1 ? S.m.y : S.m.z);
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
    // Returns the number of local nodes; sc_.nodeLocal holds the local ids.
    int partition_items(const std::vector<const int*>& rows, const std::vector<int>& lens, std::vector<int>& partOfLocal) {
        ++sc_.stamp;
        std::vector<int> ids;
        for (size_t r = 0; r < rows.size(); ++r)
            for (int k = 0; k < lens[r]; ++k) {
                const int n = rows[r][k];
                if (sc_.nodeStamp[n] != sc_.stamp) {
                    sc_.nodeStamp[n] = sc_.stamp;
                    sc_.nodeLocal[n] = (int)ids.size();
                    ids.push_back(n);
                }
            }
        partOfLocal.assign(ids.size(), 0);
        rcb(ids, 0, (int)ids.size(), 0, S.nbParts, partOfLocal);
        return (int)ids.size();
    }

    // TDTree_create_itemPart (TDTree_creation.cpp:58-85): an item whose nodes
    // all share a part joins it, any other item is a separator item.
    int item_part(const int* row, int len, const std::vector<int>& partOfLocal) {
        int a = -1, b = -1;
        for (int k = 0; k < len; ++k) {
            const int p = partOfLocal[sc_.nodeLocal[row[k]]];
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
        ++sc_.stamp;
        std::vector<int>& pos = sc_.pos;
        if ((int)pos.size() < (int)itemRev.size()) pos.assign(itemRev.size(), -1), sc_.posStamp.assign(itemRev.size(), -1);
        for (int k = fi; k <= li; ++k) {
            pos[itemRev[k]] = k - fi;
            sc_.posStamp[itemRev[k]] = sc_.stamp;
        }
        const int nCell = (int)cellPart.size();
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] == ISO) continue;
            const int c = S.t.cellRev[fc + i];
            for (int k = c2i.off[c]; k < c2i.off[c + 1]; ++k) {
                const int it = c2i.val[k];
                if (sc_.posStamp[it] != sc_.stamp) continue;
                itemPart[pos[it]] = cellPart[i];
            }
        }
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] != ISO) continue;
            const int c = S.t.cellRev[fc + i];
            for (int k = c2i.off[c]; k < c2i.off[c + 1]; ++k) {
                const int it = c2i.val[k];
                if (sc_.posStamp[it] != sc_.stamp) continue;
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

    // TDTree::TDTree_create_boundary (TDTree_creation.cpp:131-262)
    void create_boundary(int tree, int ff, int lf, int fn, int ln, bool isIso) {
        const int nFace = lf - ff + 1;
        if (nFace < S.partSize) {
            init_node(tree, 0, -1, ff, lf, fn, ln, isIso);
            return;
        }
        std::vector<const int*> rows(nFace);
        std::vector<int> lens(nFace);
        for (int i = 0; i < nFace; ++i) {
            const int f = S.t.faceRev[ff + i];
            rows[i] = S.m.f2n.data() + S.f2nOff[f];
            lens[i] = S.f2nOff[f + 1] - S.f2nOff[f];
        }
        std::vector<int> partOfLocal;
        partition_items(rows, lens, partOfLocal);
        const int parts = S.nbParts;
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
        ++sc_.stamp;
        for (int i = 0; i < nFace; ++i) {
            if (facePart[i] == ISO) continue;
            const int c1 = S.m.f2c[2 * S.t.faceRev[ff + i]];
            if (sc_.markStamp[c1] != sc_.stamp || sc_.markPart[c1] == facePart[i]) {
                sc_.markStamp[c1] = sc_.stamp;
                sc_.markPart[c1] = facePart[i];
            } else {
                --nbPartFace[facePart[i]];
                facePart[i] = ISO;
                ++nbIsoFace;
            }
        }
        std::vector<int> nbPartNode(parts, 0);
        int nbIsoNode = 0;
        if (ln - fn + 1 > 0) throw std::logic_error("boundary face tree with a node range");
        stable_by_part(S.t.faceRev, ff, facePart);

        const int nbIsoItem = std::max(nbIsoFace, nbIsoNode);
        std::vector<int> nbPartItem(parts);
        for (int i = 0; i < parts; ++i) nbPartItem[i] = std::max(nbPartFace[i], nbPartNode[i]);
        init_node(tree, 0, -1, ff, lf, fn, ln, isIso);
        (*N)[tree].nbParts = parts;
        (*N)[tree].son.assign(parts, -1);
        for (int i = 0; i < parts; ++i)
            if (nbPartItem[i] > 0) {
                int s = new_node();
                (*N)[tree].son[i] = s;
            }
        if (nbIsoItem > 0) {
            int s = new_node();
            (*N)[tree].iso = s;
        }
        bool leaf = true;
        int st2 = ff;
        for (int i = 0; i < parts; ++i) {
            const int ed2 = st2 + nbPartFace[i] - 1;
            if (nbPartItem[i] > 0) {
                leaf = false;
                create_boundary((*N)[tree].son[i], st2, ed2, 0, -1, isIso);
            }
            st2 = ed2 + 1;
        }
        (*N)[tree].isLeaf = leaf;
        if (nbIsoItem > 0 && !leaf) create_boundary((*N)[tree].iso, lf - nbIsoFace + 1, lf, 0, -1, true);
    }

    // TDTree::TDTree_create_normal (TDTree_creation.cpp:264-405)
    void create_normal(int tree, int fc, int lc, int ff, int lf, int fn, int ln, bool isIso) {
        const int nCell = lc - fc + 1;
        if (nCell < S.partSize) {
            init_node(tree, fc, lc, ff, lf, fn, ln, isIso);
            return;
        }
        std::vector<const int*> rows(nCell);
        std::vector<int> lens(nCell);
        for (int i = 0; i < nCell; ++i) {
            const int c = S.t.cellRev[fc + i];
            rows[i] = S.c2n.row(c);
            lens[i] = S.c2n.row_size(c);
        }
        std::vector<int> partOfLocal;
        partition_items(rows, lens, partOfLocal);
        const int parts = S.nbParts;
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
        ++sc_.stamp;
        for (int i = 0; i < nCell; ++i) {
            if (cellPart[i] == ISO) continue;
            const int c = S.t.cellRev[fc + i];
            const int* nb = S.c2c.row(c);
            const int nn = S.c2c.row_size(c);
            for (int j = 0; j < nn; ++j) {
                const int x = nb[j];
                if (sc_.markStamp[x] == sc_.stamp && sc_.markPart[x] != cellPart[i]) {
                    --nbPartCell[cellPart[i]];
                    ++nbIsoCell;
                    cellPart[i] = ISO;
                    break;
                }
            }
            if (cellPart[i] == ISO) continue;
            for (int j = 0; j < nn; ++j) {
                sc_.markStamp[nb[j]] = sc_.stamp;
                sc_.markPart[nb[j]] = cellPart[i];
            }
        }
        std::vector<int> facePart, nodePart, nbPartFace(parts, 0), nbPartNode(parts, 0);
        int nbIsoFace = 0, nbIsoNode = 0;
        items_accordingly(S.c2f, S.t.faceRev, fc, cellPart, ff, lf, facePart, nbPartFace, nbIsoFace);
        if (lf - ff + 1 > 0) stable_by_part(S.t.faceRev, ff, facePart);
        items_accordingly(S.c2n, S.t.nodeRev, fc, cellPart, fn, ln, nodePart, nbPartNode, nbIsoNode);
        if (ln - fn + 1 > 0) stable_by_part(S.t.nodeRev, fn, nodePart);
        stable_by_part(S.t.cellRev, fc, cellPart);

        const int nbIsoItem = std::max(nbIsoCell, std::max(nbIsoFace, nbIsoNode));
        std::vector<int> nbPartItem(parts);
        for (int i = 0; i < parts; ++i) nbPartItem[i] = std::max(nbPartCell[i], std::max(nbPartFace[i], nbPartNode[i]));
        init_node(tree, fc, lc, ff, lf, fn, ln, isIso);
        (*N)[tree].nbParts = parts;
        (*N)[tree].son.assign(parts, -1);
        for (int i = 0; i < parts; ++i)
            if (nbPartItem[i] > 0) {
                int s = new_node();
                (*N)[tree].son[i] = s;
            }
        if (nbIsoItem > 0) {
            int s = new_node();
            (*N)[tree].iso = s;
        }
        // The parts' subtrees are independent: disjoint cell, face and node
        // ranges, and no face or node shared between two parts (a cell
        // touching another part's node is a separator cell).  Large ones are
        // built side by side, each into its own node list whose entry 0 is
        // the part's node; the lists are spliced in part order afterwards,
        // which gives every node the index the serial recursion gives it.
        struct Job {
            int i, st1, ed1, st2, ed2, st3, ed3;
            std::vector<TDTreeNode> nodes;
        };
        std::vector<Job> jobs;
        {
            int st1 = fc, st2 = ff, st3 = fn;
            for (int i = 0; i < parts; ++i) {
                const int ed1 = st1 + nbPartCell[i] - 1;
                const int ed2 = st2 + nbPartFace[i] - 1;
                const int ed3 = st3 + nbPartNode[i] - 1;
                if (nbPartItem[i] > 0) jobs.push_back({i, st1, ed1, st2, ed2, st3, ed3, {}});
                st1 = ed1 + 1;
                st2 = ed2 + 1;
                st3 = ed3 + 1;
            }
        }
        const bool leaf = jobs.empty();
        if (nCell >= kParallelCells && jobs.size() > 1) {
            std::vector<std::thread> th(jobs.size());
            std::vector<std::exception_ptr> err(jobs.size());
            std::vector<TDTreeNode>* const mine = N;
            for (size_t j = 0; j < jobs.size(); ++j) {
                Job& J = jobs[j];
                J.nodes.emplace_back();
                if (j + 1 < jobs.size() && S.take_thread()) {
                    th[j] = std::thread([this, &J, &err, j, isIso] {
                        try {
                            Worker w(S, &J.nodes);
                            w.create_normal(0, J.st1, J.ed1, J.st2, J.ed2, J.st3, J.ed3, isIso);
                        } catch (...) {
                            err[j] = std::current_exception();
                        }
                        S.spare.fetch_add(1);
                    });
                } else {
                    N = &J.nodes;
                    create_normal(0, J.st1, J.ed1, J.st2, J.ed2, J.st3, J.ed3, isIso);
                    N = mine;
                }
            }
            for (size_t j = 0; j < jobs.size(); ++j)
                if (th[j].joinable()) th[j].join();
            for (const std::exception_ptr& e : err)
                if (e) std::rethrow_exception(e);
            for (Job& J : jobs) {
                const int g = (*N)[tree].son[J.i];
                const int base = (int)N->size();
                auto map = [g, base](int k) { return k < 0 ? k : (k == 0 ? g : base + k - 1); };
                for (size_t k = 0; k < J.nodes.size(); ++k) {
                    TDTreeNode d = std::move(J.nodes[k]);
                    for (int& x : d.son) x = map(x);
                    d.iso = map(d.iso);
                    if (k == 0)
                        (*N)[g] = std::move(d);
                    else
                        N->push_back(std::move(d));
                }
            }
        } else {
            for (const Job& J : jobs) create_normal((*N)[tree].son[J.i], J.st1, J.ed1, J.st2, J.ed2, J.st3, J.ed3, isIso);
        }
        (*N)[tree].isLeaf = leaf;
        if (nbIsoItem > 0 && !leaf)
            create_normal((*N)[tree].iso, lc - nbIsoCell + 1, lc, lf - nbIsoFace + 1, lf, ln - nbIsoNode + 1, ln, true);
        else
            (*N)[tree].nbParts = 0;
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
    Shared s(m, nbParts, partSize);
    std::vector<TDTreeNode> nodes;
    Worker(s, &nodes).run_root();
    TDTree& t = s.t;
    t.nodes = std::move(nodes);
    for (int i = 0; i < m.nC; ++i) t.cellPerm[t.cellRev[i]] = i;
    for (int i = 0; i < m.nF; ++i) t.facePerm[t.faceRev[i]] = i;
    for (int i = 0; i < m.nN; ++i) t.nodePerm[t.nodeRev[i]] = i;
    return std::move(t);
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
