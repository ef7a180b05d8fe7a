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

RawMesh generate_wedge(const WedgeParams &p) {
    if (p.nz != 0) throw std::runtime_error("generate_wedge: 3D meshes are built by generate_wedge_3d");
    const int nx = p.nx, ny = p.ny;
    if (nx < 2 || ny < 1) throw std::runtime_error("generate_wedge: need nx >= 2, ny >= 1");
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
    const int i_rs = col_of(rs), i_re = col_of(re);
    RawMesh m;
    m.dim = 2;
    // ---- vertices, column-major (i * (ny + 1) + j) ----
    const int64_t NV = (int64_t)(nx + 1) * (ny + 1);
    m.xyz.assign(3 * (size_t)NV, 0.0);
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
            m.xyz[3 * (size_t)v] = x;
            m.xyz[3 * (size_t)v + 1] = y;
        }
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

// Entry point used by the C-ABI: 2D generator, or the 3D extrusion.
RawMesh generate_wedge_3d(const WedgeParams &p);

RawMesh generate_wedge_any(const WedgeParams &p) {
    if (p.nz == 0) return generate_wedge(p);
    return generate_wedge_3d(p);
}

}  // namespace psst
