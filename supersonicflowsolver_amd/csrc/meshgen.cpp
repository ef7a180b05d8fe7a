// meshgen.cpp -- compression-ramp ("wedge") channel meshes in the layout of
// the reference's wedge meshes (wedge40_120.msh / wedge80_240.msh):
//   zone 2 fluid cells; boundary zones 3 symm (bottom, x < ramp start),
//   4 wedge (ramp), 5 wall (bottom after the ramp), 6 out (x = length),
//   7 up (y = height), 8 in (x = 0); zone 10 interior faces.
//   Cells are numbered column-major (ix * ny + iy) unless renumbered; for
//   every face the cell on the left of v0 -> v1 is cell(0), so the unit
//   normal (y0 - y1, x1 - x0)/A points toward cell(0)
//   (psstFiniteVolumeGridTools.cpp:408-425).  As in the reference meshes,
//   the symm zone keeps its ghost on the cell(1) side and all other boundary
//   zones on the cell(0) side, so both ghost placements are exercised.
// Random numbers: splitmix64-seeded xoshiro256** (the same generator is used
// by the Python side for parameter sweeps).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <stdexcept>

#include "psst_host.hpp"

namespace psst {

namespace {

struct Xoshiro {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double symmetric() { return 2.0 * uniform() - 1.0; }                              // [-1, 1)
};

// Reverse Cuthill-McKee order of a graph given as CSR adjacency.
std::vector<int64_t> rcm_order(const std::vector<int64_t> &off, const std::vector<int64_t> &adj) {
    const int64_t n = (int64_t)off.size() - 1;
    std::vector<int64_t> deg((size_t)n);
    for (int64_t i = 0; i < n; ++i) deg[(size_t)i] = off[(size_t)i + 1] - off[(size_t)i];
    std::vector<char> seen((size_t)n, 0);
    std::vector<int64_t> order;
    order.reserve((size_t)n);
    std::vector<int64_t> nb;
    for (int64_t start_scan = 0; start_scan < n; ++start_scan) {
        if (seen[(size_t)start_scan]) continue;
        // start from a minimum-degree vertex of this component
        int64_t start = start_scan;
        std::queue<int64_t> q;
        q.push(start);
        seen[(size_t)start] = 1;
        while (!q.empty()) {
            int64_t v = q.front();
            q.pop();
            order.push_back(v);
            nb.clear();
            for (int64_t k = off[(size_t)v]; k < off[(size_t)v + 1]; ++k)
                if (!seen[(size_t)adj[(size_t)k]]) nb.push_back(adj[(size_t)k]);
            std::sort(nb.begin(), nb.end(), [&](int64_t a, int64_t b) {
                return deg[(size_t)a] != deg[(size_t)b] ? deg[(size_t)a] < deg[(size_t)b] : a < b;
            });
            for (int64_t w : nb) {
                seen[(size_t)w] = 1;
                q.push(w);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;  // order[k] = old index placed at position k
}

}  // namespace

// Nodes of the 2D channel profile, column-major (i * (ny + 1) + j), with the
// seeded jitter: interior nodes move +-jitter of the local spacing, wall
// nodes slide along their wall, the ramp's kink nodes stay.  xy: 2 per node.
void wedge_profile(const WedgeParams &p, std::vector<double> &xy, int &i_rs, int &i_re) {
    const int nx = p.nx, ny = p.ny;
    if (nx < 2 || ny < 1) throw std::runtime_error("wedge generator: need nx >= 2, ny >= 1");
    const double L = p.length, H = p.height;
    const double dx = L / nx;
    const double tn = std::tan(p.ramp_deg * M_PI / 180.0);
    const double rs = p.ramp_start, re = p.ramp_start + p.ramp_length;
    auto ybot = [&](double x) {
        if (x <= rs) return 0.0;
        if (x >= re) return (re - rs) * tn;
        return (x - rs) * tn;
    };
    // column index of the ramp start / end (kink nodes stay fixed)
    auto col_of = [&](double x) { return (int)std::lround(x / dx); };
    i_rs = col_of(rs);
    i_re = col_of(re);
    const int64_t NV = (int64_t)(nx + 1) * (ny + 1);
    xy.assign(2 * (size_t)NV, 0.0);
    Xoshiro rng(p.jitter_seed);
    for (int i = 0; i <= nx; ++i) {
        double x0 = (i == nx) ? L : i * dx;
        if (i == i_rs) x0 = rs;
        if (i == i_re) x0 = re;
        double yb0 = ybot(x0);
        double dy = (H - yb0) / ny;
        for (int j = 0; j <= ny; ++j) {
            double x = x0;
            double y = (j == ny) ? H : yb0 + j * dy;
            double rx = rng.symmetric(), ry = rng.symmetric();
            if (p.jitter > 0.0) {
                bool left = (i == 0), right = (i == nx), bottom = (j == 0), top = (j == ny);
                bool kink = (i == i_rs || i == i_re);
                if (!left && !right && !bottom && !top) {
                    x += p.jitter * dx * rx;
                    y += p.jitter * dy * ry;
                } else if ((bottom || top) && !left && !right && !kink) {
                    x += p.jitter * dx * rx;  // slide along the wall
                    y = bottom ? ybot(x) : H;
                } else if ((left || right) && !bottom && !top) {
                    y += p.jitter * dy * ry;
                }
            }
            int64_t v = (int64_t)i * (ny + 1) + j;
            xy[2 * (size_t)v] = x;
            xy[2 * (size_t)v + 1] = y;
        }
    }
}

RawMesh generate_wedge(const WedgeParams &p) {
    if (p.nz != 0) throw std::runtime_error("generate_wedge: 3D meshes are built by generate_wedge_3d");
    const int nx = p.nx, ny = p.ny;
    int i_rs = 0, i_re = 0;
    std::vector<double> xy;
    wedge_profile(p, xy, i_rs, i_re);
    RawMesh m;
    m.dim = 2;
    // ---- vertices, column-major (i * (ny + 1) + j) ----
    const int64_t NV = (int64_t)(nx + 1) * (ny + 1);
    m.xyz.assign(3 * (size_t)NV, 0.0);
    for (int64_t v = 0; v < NV; ++v) {
        m.xyz[3 * (size_t)v] = xy[2 * (size_t)v];
        m.xyz[3 * (size_t)v + 1] = xy[2 * (size_t)v + 1];
    }
    auto vid = [&](int i, int j) { return (int64_t)i * (ny + 1) + j + 1; };
    // ---- cell numbering ----
    const int64_t NC = (int64_t)nx * ny;
    std::vector<int64_t> newid((size_t)NC);  // column-major index -> 1-based file index
    std::iota(newid.begin(), newid.end(), 1);
    if (p.numbering == 1 || p.numbering == 2) {
        std::vector<int64_t> perm((size_t)NC);  // perm[k] = column-major cell placed at position k
        if (p.numbering == 1) {
            std::vector<int64_t> off((size_t)NC + 1, 0), adj;
            adj.reserve(4 * (size_t)NC);
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j) {
                    int64_t c = (int64_t)i * ny + j;
                    if (i > 0) adj.push_back(c - ny);
                    if (j > 0) adj.push_back(c - 1);
                    if (j < ny - 1) adj.push_back(c + 1);
                    if (i < nx - 1) adj.push_back(c + ny);
                    off[(size_t)c + 1] = (int64_t)adj.size();
                }
            perm = rcm_order(off, adj);
        } else {
            std::iota(perm.begin(), perm.end(), 0);
            Xoshiro pr(p.perm_seed);
            for (int64_t k = NC - 1; k > 0; --k) {
                int64_t r = (int64_t)(pr.next() % (uint64_t)(k + 1));
                std::swap(perm[(size_t)k], perm[(size_t)r]);
            }
        }
        for (int64_t k = 0; k < NC; ++k) newid[(size_t)perm[(size_t)k]] = k + 1;
    }
    auto cid = [&](int i, int j) { return newid[(size_t)((int64_t)i * ny + j)]; };
    m.cell_zone.assign((size_t)NC, 2);
    m.cell_type.assign((size_t)NC, 3);  // quadrilaterals
    auto add_face = [&](int64_t v0, int64_t v1, int64_t c0, int64_t c1, int64_t zone) {
        m.face_nv.push_back(2);
        int64_t v[4] = {v0, v1, 0, 0};
        m.face_v.insert(m.face_v.end(), v, v + 4);
        m.face_c.push_back(c0);
        m.face_c.push_back(c1);
        m.face_zone.push_back(zone);
    };
    auto push_zone = [&](int64_t id, int64_t bc, const char *kind, const char *name, int64_t first) {
        RawZone z;
        z.id = id;
        z.bc = bc;
        z.kind = kind;
        z.name = name;
        z.first = first;
        z.last = m.n_faces();
        m.zones.push_back(z);
    };
    {
        RawZone z;
        z.id = 2; z.bc = 1; z.kind = "fluid"; z.name = "fluid"; z.first = 1; z.last = NC;
        m.zones.push_back(z);
    }
    // symm: bottom, x in [0, ramp start]; ghost on the cell(1) side
    int64_t first = m.n_faces() + 1;
    for (int i = 0; i < i_rs; ++i) add_face(vid(i, 0), vid(i + 1, 0), cid(i, 0), 0, 3);
    push_zone(3, 3, "wall", "symm", first);
    // wedge: the ramp; ghost on the cell(0) side
    first = m.n_faces() + 1;
    for (int i = i_rs; i < i_re; ++i) add_face(vid(i + 1, 0), vid(i, 0), 0, cid(i, 0), 4);
    push_zone(4, 3, "wall", "wedge", first);
    // wall: bottom after the ramp
    first = m.n_faces() + 1;
    for (int i = i_re; i < nx; ++i) add_face(vid(i + 1, 0), vid(i, 0), 0, cid(i, 0), 5);
    push_zone(5, 3, "wall", "wall", first);
    // out: x = length
    first = m.n_faces() + 1;
    for (int j = 0; j < ny; ++j) add_face(vid(nx, j + 1), vid(nx, j), 0, cid(nx - 1, j), 6);
    push_zone(6, 3, "wall", "out", first);
    // up: y = height
    first = m.n_faces() + 1;
    for (int i = 0; i < nx; ++i) add_face(vid(i, ny), vid(i + 1, ny), 0, cid(i, ny - 1), 7);
    push_zone(7, 3, "wall", "up", first);
    // in: x = 0
    first = m.n_faces() + 1;
    for (int j = 0; j < ny; ++j) add_face(vid(0, j), vid(0, j + 1), 0, cid(0, j), 8);
    push_zone(8, 3, "wall", "in", first);
    // interior faces, ordered by their lower-numbered cell
    struct IF {
        int64_t key, v0, v1, c0, c1;
    };
    std::vector<IF> inner;
    inner.reserve(2 * (size_t)NC);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            if (i < nx - 1) {  // vertical face x_{i+1}: c0 = left cell, c1 = right cell, v0 -> v1 upward
                int64_t a = cid(i, j), b = cid(i + 1, j);
                inner.push_back({std::min(a, b) * 2, vid(i + 1, j), vid(i + 1, j + 1), a, b});
            }
            if (j < ny - 1) {  // horizontal face y_{j+1}: c0 = upper cell, c1 = lower cell, v0 -> v1 along +x
                int64_t a = cid(i, j + 1), b = cid(i, j);
                inner.push_back({std::min(a, b) * 2 + 1, vid(i, j + 1), vid(i + 1, j + 1), a, b});
            }
        }
    if (p.numbering != 0)
        std::stable_sort(inner.begin(), inner.end(), [](const IF &a, const IF &b) { return a.key < b.key; });
    first = m.n_faces() + 1;
    for (const IF &f : inner) add_face(f.v0, f.v1, f.c0, f.c1, 10);
    push_zone(10, 2, "interior", "default-interior", first);
    return m;
}

}  // namespace psst

namespace psst {

namespace {

// Hexahedron order of the extrusion: column-major (i, j, k; k fastest), RCM
// over the 6-neighbour graph, or a seeded random permutation.
std::vector<int64_t> hex_numbering(const WedgeParams &p) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int64_t NH = (int64_t)nx * ny * nz;
    std::vector<int64_t> newid((size_t)NH);  // column-major hex -> 0-based position
    std::iota(newid.begin(), newid.end(), 0);
    if (p.numbering == 0) return newid;
    std::vector<int64_t> perm((size_t)NH);
    if (p.numbering == 1) {
        auto h = [&](int i, int j, int k) { return ((int64_t)i * ny + j) * nz + k; };
        std::vector<int64_t> off((size_t)NH + 1, 0), adj;
        adj.reserve(6 * (size_t)NH);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    if (i > 0) adj.push_back(h(i - 1, j, k));
                    if (j > 0) adj.push_back(h(i, j - 1, k));
                    if (k > 0) adj.push_back(h(i, j, k - 1));
                    if (k < nz - 1) adj.push_back(h(i, j, k + 1));
                    if (j < ny - 1) adj.push_back(h(i, j + 1, k));
                    if (i < nx - 1) adj.push_back(h(i + 1, j, k));
                    off[(size_t)h(i, j, k) + 1] = (int64_t)adj.size();
                }
        perm = rcm_order(off, adj);
    } else {
        std::iota(perm.begin(), perm.end(), 0);
        Xoshiro pr(p.perm_seed);
        for (int64_t k = NH - 1; k > 0; --k) {
            int64_t r = (int64_t)(pr.next() % (uint64_t)(k + 1));
            std::swap(perm[(size_t)k], perm[(size_t)r]);
        }
    }
    for (int64_t k = 0; k < NH; ++k) newid[(size_t)perm[(size_t)k]] = k;
    return newid;
}

// The six tetrahedra of a hexahedron (Kuhn split along its 000 -> 111
// diagonal): tetrahedron sigma = (a, b, c), a permutation of the axes, has
// the corners 000, e_a, e_a + e_b, 111.  Every hexahedron is split the same
// way, so the triangles of neighbouring hexahedra match.
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
int perm_index(int a, int b, int c) {
    for (int t = 0; t < 6; ++t)
        if (kPerm[t][0] == a && kPerm[t][1] == b && kPerm[t][2] == c) return t;
    return -1;
}

}  // namespace

// Extrusion of the 2D channel profile over nz layers of `depth` along z:
// hexahedra, or six tetrahedra per hexahedron.  Zones as in 2D (symm keeps
// its ghost on the cell(1) side) plus zone 9 for the z = 0 and z = depth
// faces: slip walls "frontAndBack", or, with periodic_z, one zone of Fluent
// type 12 ("periodic") whose faces join the cells of the first and the last
// layer (faces on the z = 0 plane; cell(0) in the first layer).  Face
// vertices are ordered so that the normal that the reference computes
// (psstFiniteVolumeGridTools.cpp:426-501, right-hand rule) points into
// cell(0).
RawMesh generate_wedge_3d(const WedgeParams &p) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    if (nz < 1) throw std::runtime_error("generate_wedge_3d: nz >= 1");
    if (p.periodic_z && nz < 2) throw std::runtime_error("generate_wedge_3d: a periodic extrusion needs nz >= 2");
    int i_rs = 0, i_re = 0;
    std::vector<double> xy;
    wedge_profile(p, xy, i_rs, i_re);
    RawMesh m;
    m.dim = 3;
    const int64_t NV = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    m.xyz.assign(3 * (size_t)NV, 0.0);
    auto vid0 = [&](int i, int j, int k) { return ((int64_t)i * (ny + 1) + j) * (nz + 1) + k; };
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                const int64_t v = vid0(i, j, k);
                const int64_t q = (int64_t)i * (ny + 1) + j;
                m.xyz[3 * (size_t)v] = xy[2 * (size_t)q];
                m.xyz[3 * (size_t)v + 1] = xy[2 * (size_t)q + 1];
                m.xyz[3 * (size_t)v + 2] = (k == nz) ? p.depth : p.depth * k / nz;
            }
    const std::vector<int64_t> hpos = hex_numbering(p);
    const bool tets = p.tets != 0;
    const int64_t NH = (int64_t)nx * ny * nz, NC = tets ? 6 * NH : NH;
    m.cell_zone.assign((size_t)NC, 2);
    m.cell_type.assign((size_t)NC, tets ? 2 : 4);  // tetrahedra or hexahedra
    auto hcm = [&](int i, int j, int k) { return ((int64_t)i * ny + j) * nz + k; };
    // 1-based file id of a hex, or of tetrahedron sigma of a hex
    auto hex_cell = [&](int i, int j, int k) { return hpos[(size_t)hcm(i, j, k)] + 1; };
    auto tet_cell = [&](int i, int j, int k, int a, int b, int c) {
        return hpos[(size_t)hcm(i, j, k)] * 6 + perm_index(a, b, c) + 1;
    };
    // corner (d = 0/1 per axis) of hex (i, j, k), 1-based vertex id
    auto corner = [&](int i, int j, int k, int dx, int dy, int dz) { return vid0(i + dx, j + dy, k + dz) + 1; };
    auto X = [&](int64_t v1, int c) { return m.xyz[3 * (size_t)(v1 - 1) + c]; };
    // approximate centre of a cell from the vertices that define it
    auto centre_of = [&](const int64_t *v, int n, double *out) {
        out[0] = out[1] = out[2] = 0.0;
        for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) out[c] += X(v[q], c);
        for (int c = 0; c < 3; ++c) out[c] /= n;
    };
    auto hex_centre = [&](int i, int j, int k, double *out) {
        int64_t v[8];
        int q = 0;
        for (int dx = 0; dx < 2; ++dx)
            for (int dy = 0; dy < 2; ++dy)
                for (int dz = 0; dz < 2; ++dz) v[q++] = corner(i, j, k, dx, dy, dz);
        centre_of(v, 8, out);
    };
    auto tet_verts = [&](int i, int j, int k, int a, int b, int c, int64_t *v) {
        (void)c;
        int d[3] = {0, 0, 0};
        v[0] = corner(i, j, k, 0, 0, 0);
        d[a] = 1;
        v[1] = corner(i, j, k, d[0], d[1], d[2]);
        d[b] = 1;
        v[2] = corner(i, j, k, d[0], d[1], d[2]);
        v[3] = corner(i, j, k, 1, 1, 1);
    };
    // adds a face; `toward` is a point on the cell(0) side (its centre), or,
    // when cell(0) is a ghost, a point inside cell(1) with flip = true
    auto add_face = [&](int nv, const int64_t *vin, int64_t c0, int64_t c1, int64_t zone, const double *toward,
                        bool flip) {
        int64_t v[4] = {vin[0], vin[1], vin[2], nv == 4 ? vin[3] : 0};
        double avg[3] = {0, 0, 0};
        for (int q = 0; q < nv; ++q)
            for (int c = 0; c < 3; ++c) avg[c] += X(v[q], c) / nv;
        double a[3], b[3];
        for (int c = 0; c < 3; ++c) {
            a[c] = X(v[nv - 1], c) - avg[c];
            b[c] = X(v[0], c) - avg[c];
        }
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        double dot = 0.0;
        for (int c = 0; c < 3; ++c) dot += n[c] * (toward[c] - avg[c]);
        if (flip) dot = -dot;
        if (dot < 0.0) std::reverse(v, v + nv);
        m.face_nv.push_back((int8_t)nv);
        m.face_v.insert(m.face_v.end(), v, v + 4);
        m.face_c.push_back(c0);
        m.face_c.push_back(c1);
        m.face_zone.push_back(zone);
    };
    auto push_zone = [&](int64_t id, int64_t bc, const char *kind, const char *name, int64_t first) {
        RawZone z;
        z.id = id;
        z.bc = bc;
        z.kind = kind;
        z.name = name;
        z.first = first;
        z.last = m.n_faces();
        m.zones.push_back(z);
    };
    {
        RawZone z;
        z.id = 2;
        z.bc = 1;
        z.kind = "fluid";
        z.name = "fluid";
        z.first = 1;
        z.last = NC;
        m.zones.push_back(z);
    }
    // Boundary face on side s (0: minus, 1: plus) of axis `ax` of hex (i,j,k),
    // whose only cell is that hex (or its two tetrahedra on that side).
    // ghost_c0: the ghost is cell(0) (normal points out of the domain).
    auto boundary_face = [&](int i, int j, int k, int ax, int s, int64_t zone, bool ghost_c0) {
        double cc[3];
        hex_centre(i, j, k, cc);
        const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;  // the two other axes
        if (!tets) {
            int64_t v[4];
            int q = 0;
            int d[3];
            d[ax] = s;
            const int loop[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int t = 0; t < 4; ++t) {
                d[b] = loop[t][0];
                d[c] = loop[t][1];
                v[q++] = corner(i, j, k, d[0], d[1], d[2]);
            }
            const int64_t h = hex_cell(i, j, k);
            add_face(4, v, ghost_c0 ? 0 : h, ghost_c0 ? h : 0, zone, cc, ghost_c0);
            return;
        }
        // two triangles; plus side: tets (ax,b,c), (ax,c,b); minus side: (b,c,ax), (c,b,ax)
        for (int t = 0; t < 2; ++t) {
            const int o1 = t == 0 ? b : c, o2 = t == 0 ? c : b;
            int64_t tv[4], v[3];
            int sa, sb, sc;
            if (s == 1) {
                sa = ax; sb = o1; sc = o2;
                tet_verts(i, j, k, sa, sb, sc, tv);
                v[0] = tv[1]; v[1] = tv[2]; v[2] = tv[3];  // e_a, e_a+e_b, 111
            } else {
                sa = o1; sb = o2; sc = ax;
                tet_verts(i, j, k, sa, sb, sc, tv);
                v[0] = tv[0]; v[1] = tv[1]; v[2] = tv[2];  // 000, e_b, e_b+e_c
            }
            const int64_t cell = tet_cell(i, j, k, sa, sb, sc);
            double tc[3];
            centre_of(tv, 4, tc);
            add_face(3, v, ghost_c0 ? 0 : cell, ghost_c0 ? cell : 0, zone, tc, ghost_c0);
        }
    };
    int64_t first = m.n_faces() + 1;
    for (int i = 0; i < i_rs; ++i)
        for (int k = 0; k < nz; ++k) boundary_face(i, 0, k, 1, 0, 3, false);
    push_zone(3, 3, "wall", "symm", first);
    first = m.n_faces() + 1;
    for (int i = i_rs; i < i_re; ++i)
        for (int k = 0; k < nz; ++k) boundary_face(i, 0, k, 1, 0, 4, true);
    push_zone(4, 3, "wall", "wedge", first);
    first = m.n_faces() + 1;
    for (int i = i_re; i < nx; ++i)
        for (int k = 0; k < nz; ++k) boundary_face(i, 0, k, 1, 0, 5, true);
    push_zone(5, 3, "wall", "wall", first);
    first = m.n_faces() + 1;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) boundary_face(nx - 1, j, k, 0, 1, 6, true);
    push_zone(6, 3, "wall", "out", first);
    first = m.n_faces() + 1;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) boundary_face(i, ny - 1, k, 1, 1, 7, true);
    push_zone(7, 3, "wall", "up", first);
    first = m.n_faces() + 1;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) boundary_face(0, j, k, 0, 0, 8, true);
    push_zone(8, 3, "wall", "in", first);
    // Face between hex (i,j,k) (minus side) and hex (i',j',k') (plus side)
    // along axis ax; with tetrahedra, its two triangles.
    auto inner_face = [&](int i, int j, int k, int ax, int i2, int j2, int k2, int64_t zone) {
        const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;
        if (!tets) {
            int64_t v[4];
            int d[3];
            d[ax] = 1;
            const int loop[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int t = 0; t < 4; ++t) {
                d[b] = loop[t][0];
                d[c] = loop[t][1];
                v[t] = corner(i, j, k, d[0], d[1], d[2]);
            }
            if (zone != 10) {  // periodic: the vertices of the plus-side hex's minus face
                d[ax] = 0;
                for (int t = 0; t < 4; ++t) {
                    d[b] = loop[t][0];
                    d[c] = loop[t][1];
                    v[t] = corner(i2, j2, k2, d[0], d[1], d[2]);
                }
            }
            double cc[3];
            hex_centre(i2, j2, k2, cc);  // cell(0) = the plus-side hex
            add_face(4, v, hex_cell(i2, j2, k2), hex_cell(i, j, k), zone, cc, false);
            return;
        }
        for (int t = 0; t < 2; ++t) {
            const int o1 = t == 0 ? b : c, o2 = t == 0 ? c : b;
            int64_t tv[4], v[3];
            // plus-side tetrahedron (o1, o2, ax) of the second hex: its face 000, e_o1, e_o1+e_o2
            tet_verts(i2, j2, k2, o1, o2, ax, tv);
            v[0] = tv[0];
            v[1] = tv[1];
            v[2] = tv[2];
            double tc[3];
            centre_of(tv, 4, tc);
            add_face(3, v, tet_cell(i2, j2, k2, o1, o2, ax), tet_cell(i, j, k, ax, o1, o2), zone, tc, false);
        }
    };
    first = m.n_faces() + 1;
    if (p.periodic_z) {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) inner_face(i, j, nz - 1, 2, i, j, 0, 9);
        push_zone(9, 12, "periodic", "periodic", first);
    } else {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                boundary_face(i, j, 0, 2, 0, 9, true);
                boundary_face(i, j, nz - 1, 2, 1, 9, true);
            }
        push_zone(9, 3, "wall", "frontAndBack", first);
    }
    first = m.n_faces() + 1;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (i < nx - 1) inner_face(i, j, k, 0, i + 1, j, k, 10);
                if (j < ny - 1) inner_face(i, j, k, 1, i, j + 1, k, 10);
                if (k < nz - 1) inner_face(i, j, k, 2, i, j, k + 1, 10);
                if (tets) {
                    // the six faces inside the hexahedron: (000, e_a+e_b, 111) between
                    // (a,b,c) and (b,a,c); (000, e_a, 111) between (a,b,c) and (a,c,b)
                    for (int c = 0; c < 3; ++c) {
                        const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
                        int64_t tv[4], v[3];
                        tet_verts(i, j, k, a, b, c, tv);
                        v[0] = tv[0];
                        v[1] = tv[2];
                        v[2] = tv[3];
                        double tc[3];
                        centre_of(tv, 4, tc);
                        add_face(3, v, tet_cell(i, j, k, a, b, c), tet_cell(i, j, k, b, a, c), 10, tc, false);
                    }
                    for (int a = 0; a < 3; ++a) {
                        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
                        int64_t tv[4], v[3];
                        tet_verts(i, j, k, a, b, c, tv);
                        v[0] = tv[0];
                        v[1] = tv[1];
                        v[2] = tv[3];
                        double tc[3];
                        centre_of(tv, 4, tc);
                        add_face(3, v, tet_cell(i, j, k, a, b, c), tet_cell(i, j, k, a, c, b), 10, tc, false);
                    }
                }
            }
    push_zone(10, 2, "interior", "default-interior", first);
    return m;
}

// Entry point used by the C-ABI: 2D generator, or the 3D extrusion.
RawMesh generate_wedge_any(const WedgeParams &p) {
    if (p.nz == 0) return generate_wedge(p);
    return generate_wedge_3d(p);
}

}  // namespace psst
