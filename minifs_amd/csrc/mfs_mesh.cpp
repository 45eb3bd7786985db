// Box+bump Kuhn-tetrahedron generator and FlowStar binary I/O.
//
// The reader follows src/Kernels/load_grid.cpp:129-273 (boundary faces
// reoriented when the first cell is a boundary side, nBFace counted but not
// reordered) and ReadMFlowBoundaryCondition (load_grid.cpp:282-314).
#include "mfs_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mfs {

std::vector<int> RawMesh::f2n_offsets() const {
    std::vector<int> off(nF + 1);
    off[0] = 0;
    for (int i = 0; i < nF; ++i) off[i + 1] = off[i] + nNPF[i];
    return off;
}

namespace {

struct FaceRec {
    int a, b, c;      // sorted node ids (key)
    int cell;         // owning tetrahedron
    int opp;          // node of that tetrahedron opposite the face
    int n0, n1, n2;   // node ids in tetrahedron-local order
};

inline bool key_less(const FaceRec& l, const FaceRec& r) {
    if (l.a != r.a) return l.a < r.a;
    if (l.b != r.b) return l.b < r.b;
    if (l.c != r.c) return l.c < r.c;
    return l.cell < r.cell;
}

}  // namespace

namespace {

// Hexahedra or prisms: every cell lists its faces with right-hand normals
// pointing out of the cell, faces are matched by their sorted node sets, and
// an interior face takes the node order of its lower-numbered cell (so its
// normal points from c1 to c2).
void build_poly_faces(RawMesh& m, const BoxBumpSpec& s) {
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const int px = nx + 1, py = ny + 1;
    auto nid = [&](int i, int j, int k) { return i + px * (j + py * k); };
    struct PF {
        int key[4];  // sorted node ids, -1 padded
        int cell;
        int nn;
        int n[4];    // outward order, smallest id first
    };
    std::vector<PF> recs;
    recs.reserve((size_t)nx * ny * nz * (s.cells == CELLS_HEX ? 6 : 10));
    auto add = [&](int cell, std::initializer_list<int> nodes) {
        PF r;
        r.cell = cell;
        r.nn = (int)nodes.size();
        int t[4], u = 0;
        for (int v : nodes) t[u++] = v;
        int rot = 0;
        for (int a = 1; a < r.nn; ++a)
            if (t[a] < t[rot]) rot = a;
        for (int a = 0; a < r.nn; ++a) r.n[a] = t[(rot + a) % r.nn];
        for (int a = 0; a < 4; ++a) r.key[a] = a < r.nn ? t[a] : -1;
        std::sort(r.key, r.key + r.nn);
        if (r.nn < 4) r.n[3] = -1;
        recs.push_back(r);
    };
    int cell = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int v0 = nid(i, j, k), v1 = nid(i + 1, j, k), v2 = nid(i + 1, j + 1, k), v3 = nid(i, j + 1, k);
                const int v4 = nid(i, j, k + 1), v5 = nid(i + 1, j, k + 1), v6 = nid(i + 1, j + 1, k + 1),
                          v7 = nid(i, j + 1, k + 1);
                if (s.cells == CELLS_HEX) {
                    add(cell, {v0, v3, v2, v1});
                    add(cell, {v4, v5, v6, v7});
                    add(cell, {v0, v1, v5, v4});
                    add(cell, {v3, v7, v6, v2});
                    add(cell, {v0, v4, v7, v3});
                    add(cell, {v1, v2, v6, v5});
                    ++cell;
                } else {
                    // prism A: (0,1,2) / (4,5,6); prism B: (0,2,3) / (4,6,7), bottoms counter-clockwise
                    const int pr[2][6] = {{v0, v1, v2, v4, v5, v6}, {v0, v2, v3, v4, v6, v7}};
                    for (int h = 0; h < 2; ++h) {
                        const int a = pr[h][0], b = pr[h][1], c = pr[h][2], a2 = pr[h][3], b2 = pr[h][4],
                                  c2 = pr[h][5];
                        add(cell, {a, c, b});
                        add(cell, {a2, b2, c2});
                        add(cell, {a, b, b2, a2});
                        add(cell, {b, c, c2, b2});
                        add(cell, {c, a, a2, c2});
                        ++cell;
                    }
                }
            }
    m.nC = cell;
    std::sort(recs.begin(), recs.end(), [](const PF& l, const PF& r) {
        for (int a = 0; a < 4; ++a)
            if (l.key[a] != r.key[a]) return l.key[a] < r.key[a];
        return l.cell < r.cell;
    });
    struct Face {
        int c1, c2, nn;
        int n[4];
    };
    std::vector<Face> bf, inf;
    auto same = [](const PF& l, const PF& r) {
        for (int a = 0; a < 4; ++a)
            if (l.key[a] != r.key[a]) return false;
        return true;
    };
    for (size_t p = 0; p < recs.size();) {
        size_t q = p + 1;
        while (q < recs.size() && same(recs[q], recs[p])) ++q;
        const PF& lo = recs[p];  // lower cell first (sort tie-break)
        Face f;
        f.c1 = lo.cell;
        f.nn = lo.nn;
        for (int a = 0; a < 4; ++a) f.n[a] = lo.n[a];
        if (q - p == 1) {
            bool bottom = true;
            for (int a = 0; a < lo.nn; ++a) bottom = bottom && (lo.n[a] / (px * py) == 0);
            f.c2 = bottom ? -1 : -2;
            bf.push_back(f);
        } else {
            f.c2 = recs[p + 1].cell;
            inf.push_back(f);
        }
        p = q;
    }
    std::sort(bf.begin(), bf.end(), [](const Face& l, const Face& r) {
        if (l.c1 != r.c1) return l.c1 < r.c1;
        if (l.n[0] != r.n[0]) return l.n[0] < r.n[0];
        return l.n[1] < r.n[1];
    });
    std::sort(inf.begin(), inf.end(), [](const Face& l, const Face& r) {
        if (l.c1 != r.c1) return l.c1 < r.c1;
        return l.c2 < r.c2;
    });
    m.nB = (int)bf.size();
    m.nF = (int)(bf.size() + inf.size());
    m.nNPF.clear();
    m.f2n.clear();
    m.f2c.clear();
    for (const std::vector<Face>* list : {&bf, &inf})
        for (const Face& f : *list) {
            m.nNPF.push_back(f.nn);
            for (int a = 0; a < f.nn; ++a) m.f2n.push_back(f.n[a]);
            m.f2c.push_back(f.c1);
            m.f2c.push_back(f.c2);
        }
}

}  // namespace

RawMesh make_box_bump(const BoxBumpSpec& s) {
    RawMesh m;
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    m.nN = px * py * pz;
    m.nC = 6 * nx * ny * nz;
    m.x.resize(m.nN);
    m.y.resize(m.nN);
    m.z.resize(m.nN);
    auto nid = [&](int i, int j, int k) { return i + px * (j + py * k); };

    SplitMix64 rng(s.seed);
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    const double two_s2 = 2.0 * s.bump_sigma * s.bump_sigma;
    for (int k = 0; k < pz; ++k)
        for (int j = 0; j < py; ++j)
            for (int i = 0; i < px; ++i) {
                double xv = i * hx, yv = j * hy, zv = k * hz;
                if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
                    xv += s.jitter * hx * rng.uniform_pm1();
                    yv += s.jitter * hy * rng.uniform_pm1();
                    zv += s.jitter * hz * rng.uniform_pm1();
                }
                const double dxb = xv - 0.5, dyb = yv - 0.5;
                const double bump = s.bump_height * std::exp(-(dxb * dxb + dyb * dyb) / two_s2);
                zv = zv + bump * (1.0 - zv);
                const int n = nid(i, j, k);
                m.x[n] = xv;
                m.y[n] = yv;
                m.z[n] = zv;
            }

    if (s.cells != CELLS_TET) {
        build_poly_faces(m, s);
        m.bc_patch = {1, 2};
        m.bc_type = {BC_WALL, BC_FAR_FIELD};
        return m;
    }

    // Kuhn split: tetrahedron t of a hex follows the axis permutation perm[t]
    // from corner (0,0,0) to corner (1,1,1).
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<FaceRec> recs((size_t)m.nC * 4);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int hex = i + nx * (j + ny * k);
                for (int t = 0; t < 6; ++t) {
                    int v[4];
                    int off[3] = {0, 0, 0};
                    v[0] = nid(i, j, k);
                    for (int st = 0; st < 3; ++st) {
                        off[perm[t][st]] = 1;
                        v[st + 1] = nid(i + off[0], j + off[1], k + off[2]);
                    }
                    const int cell = 6 * hex + t;
                    for (int f = 0; f < 4; ++f) {
                        int q[3], c = 0;
                        for (int u = 0; u < 4; ++u)
                            if (u != f) q[c++] = v[u];
                        FaceRec r;
                        r.n0 = q[0];
                        r.n1 = q[1];
                        r.n2 = q[2];
                        std::sort(q, q + 3);
                        r.a = q[0];
                        r.b = q[1];
                        r.c = q[2];
                        r.cell = cell;
                        r.opp = v[f];
                        recs[(size_t)cell * 4 + f] = r;
                    }
                }
            }
    std::sort(recs.begin(), recs.end(), key_less);

    struct Face {
        int c1, c2;  // c2 < 0: boundary patch
        int n[3];
        int local;   // tie-break for boundary ordering
    };
    std::vector<Face> bfaces, ifaces;
    bfaces.reserve(2 * (nx * ny + ny * nz + nx * nz) * 2);
    ifaces.reserve((size_t)m.nC * 2);

    auto orient = [&](int c1opp, int n0, int n1, int n2, int out[3]) {
        // Right-hand normal of (n0,n1,n2) must point away from the opposite node of c1.
        const double ax = m.x[n1] - m.x[n0], ay = m.y[n1] - m.y[n0], az = m.z[n1] - m.z[n0];
        const double bx = m.x[n2] - m.x[n0], by = m.y[n2] - m.y[n0], bz = m.z[n2] - m.z[n0];
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double d = cx * (m.x[c1opp] - m.x[n0]) + cy * (m.y[c1opp] - m.y[n0]) + cz * (m.z[c1opp] - m.z[n0]);
        int t[3] = {n0, n1, n2};
        if (d > 0) std::swap(t[1], t[2]);
        // rotate so the smallest node id comes first (cyclic order kept)
        int r = 0;
        if (t[1] < t[r]) r = 1;
        if (t[2] < t[r]) r = 2;
        for (int u = 0; u < 3; ++u) out[u] = t[(r + u) % 3];
    };

    auto node_ijk = [&](int n, int& i, int& j, int& k) {
        i = n % px;
        j = (n / px) % py;
        k = n / (px * py);
    };

    size_t p = 0;
    while (p < recs.size()) {
        size_t q = p + 1;
        while (q < recs.size() && recs[q].a == recs[p].a && recs[q].b == recs[p].b && recs[q].c == recs[p].c) ++q;
        const FaceRec& r0 = recs[p];
        if (q - p == 1) {
            Face f;
            f.c1 = r0.cell;
            // patch from the box side holding all three nodes
            int i0, j0, k0, i1, j1, k1, i2, j2, k2;
            node_ijk(r0.a, i0, j0, k0);
            node_ijk(r0.b, i1, j1, k1);
            node_ijk(r0.c, i2, j2, k2);
            const bool bottom = (k0 == 0 && k1 == 0 && k2 == 0);
            f.c2 = bottom ? -1 : -2;
            orient(r0.opp, r0.n0, r0.n1, r0.n2, f.n);
            f.local = 0;
            bfaces.push_back(f);
        } else {
            const FaceRec& r1 = recs[p + 1];
            const FaceRec& lo = (r0.cell < r1.cell) ? r0 : r1;
            const FaceRec& hi = (r0.cell < r1.cell) ? r1 : r0;
            Face f;
            f.c1 = lo.cell;
            f.c2 = hi.cell;
            orient(lo.opp, lo.n0, lo.n1, lo.n2, f.n);
            f.local = 0;
            ifaces.push_back(f);
        }
        p = q;
    }
    std::vector<FaceRec>().swap(recs);
    std::sort(bfaces.begin(), bfaces.end(), [](const Face& l, const Face& r) {
        if (l.c1 != r.c1) return l.c1 < r.c1;
        if (l.n[0] != r.n[0]) return l.n[0] < r.n[0];
        return l.n[1] < r.n[1];
    });
    std::sort(ifaces.begin(), ifaces.end(), [](const Face& l, const Face& r) {
        if (l.c1 != r.c1) return l.c1 < r.c1;
        return l.c2 < r.c2;
    });

    m.nB = (int)bfaces.size();
    m.nF = (int)(bfaces.size() + ifaces.size());
    m.nNPF.assign(m.nF, 3);
    m.f2n.resize((size_t)m.nF * 3);
    m.f2c.resize((size_t)m.nF * 2);
    int fi = 0;
    for (const Face& f : bfaces) {
        m.f2n[3 * fi] = f.n[0];
        m.f2n[3 * fi + 1] = f.n[1];
        m.f2n[3 * fi + 2] = f.n[2];
        m.f2c[2 * fi] = f.c1;
        m.f2c[2 * fi + 1] = f.c2;
        ++fi;
    }
    for (const Face& f : ifaces) {
        m.f2n[3 * fi] = f.n[0];
        m.f2n[3 * fi + 1] = f.n[1];
        m.f2n[3 * fi + 2] = f.n[2];
        m.f2c[2 * fi] = f.c1;
        m.f2c[2 * fi + 1] = f.c2;
        ++fi;
    }
    m.bc_patch = {1, 2};
    m.bc_type = {BC_WALL, BC_FAR_FIELD};
    return m;
}

bool write_flowstar(const RawMesh& m, const std::string& mesh_path, const std::string& bc_path, std::string* err) {
    FILE* fp = std::fopen(mesh_path.c_str(), "wb");
    if (!fp) {
        if (err) *err = "cannot open " + mesh_path + " for writing";
        return false;
    }
    int hdr[3] = {m.nN, m.nF, m.nC};
    bool ok = std::fwrite(hdr, sizeof(int), 3, fp) == 3;
    ok = ok && std::fwrite(m.x.data(), sizeof(double), m.nN, fp) == (size_t)m.nN;
    ok = ok && std::fwrite(m.y.data(), sizeof(double), m.nN, fp) == (size_t)m.nN;
    ok = ok && std::fwrite(m.z.data(), sizeof(double), m.nN, fp) == (size_t)m.nN;
    ok = ok && std::fwrite(m.nNPF.data(), sizeof(int), m.nF, fp) == (size_t)m.nF;
    ok = ok && std::fwrite(m.f2n.data(), sizeof(int), m.f2n.size(), fp) == m.f2n.size();
    ok = ok && std::fwrite(m.f2c.data(), sizeof(int), m.f2c.size(), fp) == m.f2c.size();
    int tail[2] = {0, 0};  // nIFace, nINode: the reader always reads them (load_grid.cpp:235-241)
    ok = ok && std::fwrite(tail, sizeof(int), 2, fp) == 2;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) {
        if (err) *err = "short write to " + mesh_path;
        return false;
    }
    std::ofstream bc(bc_path);
    if (!bc) {
        if (err) *err = "cannot open " + bc_path + " for writing";
        return false;
    }
    bc << m.bc_patch.size() << "\n";
    for (size_t i = 0; i < m.bc_patch.size(); ++i) bc << m.bc_patch[i] << " " << m.bc_type[i] << "\n";
    return (bool)bc;
}

bool read_flowstar(RawMesh& m, const std::string& mesh_path, const std::string& bc_path, std::string* err) {
    FILE* fp = std::fopen(mesh_path.c_str(), "rb");
    if (!fp) {
        if (err) *err = "could not open file " + mesh_path;
        return false;
    }
    auto fail = [&](const char* what) {
        std::fclose(fp);
        if (err) *err = std::string(what) + " in " + mesh_path;
        return false;
    };
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, fp) != 3) return fail("truncated header");
    m.nN = hdr[0];
    m.nF = hdr[1];
    m.nC = hdr[2];
    if (m.nN <= 0 || m.nF <= 0 || m.nC <= 0) return fail("bad sizes");
    m.x.resize(m.nN);
    m.y.resize(m.nN);
    m.z.resize(m.nN);
    if (std::fread(m.x.data(), sizeof(double), m.nN, fp) != (size_t)m.nN) return fail("truncated x");
    if (std::fread(m.y.data(), sizeof(double), m.nN, fp) != (size_t)m.nN) return fail("truncated y");
    if (std::fread(m.z.data(), sizeof(double), m.nN, fp) != (size_t)m.nN) return fail("truncated z");
    m.nNPF.resize(m.nF);
    if (std::fread(m.nNPF.data(), sizeof(int), m.nF, fp) != (size_t)m.nF) return fail("truncated nNPF");
    std::vector<int> off = m.f2n_offsets();
    m.f2n.resize(off[m.nF]);
    if (std::fread(m.f2n.data(), sizeof(int), m.f2n.size(), fp) != m.f2n.size()) return fail("truncated f2n");
    m.f2c.resize(2 * (size_t)m.nF);
    if (std::fread(m.f2c.data(), sizeof(int), m.f2c.size(), fp) != m.f2c.size()) return fail("truncated f2c");
    std::fclose(fp);
    // load_grid.cpp:206-232: a face whose first side is the boundary gets its
    // nodes 1..n-1 reversed and its sides swapped; boundary faces are counted.
    int nB = 0;
    for (int i = 0; i < m.nF; ++i) {
        int& c1 = m.f2c[2 * i];
        int& c2 = m.f2c[2 * i + 1];
        if (c1 < 0) {
            std::reverse(m.f2n.begin() + off[i] + 1, m.f2n.begin() + off[i + 1]);
            std::swap(c1, c2);
            ++nB;
        } else if (c2 < 0) {
            ++nB;
        }
    }
    m.nB = nB;
    for (int i = 0; i < nB; ++i)
        if (m.f2c[2 * i + 1] >= 0) {
            if (err) *err = "boundary faces must be the first nBFace faces";
            return false;
        }
    std::ifstream in(bc_path);
    if (!in) {
        if (err) *err = "failed to open " + bc_path;
        return false;
    }
    int n_patch = 0;
    in >> n_patch;
    m.bc_patch.clear();
    m.bc_type.clear();
    for (int i = 0; i < n_patch; ++i) {
        int pid, type;
        if (!(in >> pid >> type)) {
            if (err) *err = "truncated " + bc_path;
            return false;
        }
        m.bc_patch.push_back(pid);
        m.bc_type.push_back(type);
    }
    return true;
}

}  // namespace mfs
