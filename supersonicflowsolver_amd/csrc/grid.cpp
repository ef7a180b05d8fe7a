// grid.cpp -- flattening of a Fluent mesh into the finite-volume grid the
// device consumes: canonical numbering, FP64 geometry and the reference's
// iteration orders.
//
// What is kept from the reference (psstAnsysInputGrid::create_finite_volume_grid,
// psstAnsysGridReader.cpp:688-907) and why:
//   * one ghost cell per boundary face, its centre = the face centre, V = 0
//     (:770-781, psstFiniteVolumeGridTools.cpp:725-729);
//   * adapted meshes: parent cells (zone type 32) and parent faces (type 31)
//     take no part in the solve; a parent face joins no cell's face list and
//     has no ghost (:860-874), and the flux and gradient loops skip it;
//   * a cell's face list is built by visiting face zones in FILE order, then
//     faces in zone order (:825-883).  The time-step and least-squares sums run
//     over this list, so its order decides their rounding;
//   * the flux and Green-Gauss sums run over faces in ascending zone id order
//     (std::map<zone id, psstFaceStorage>), which is the canonical face order;
//   * geometry formulas and their operation order
//     (psstFiniteVolumeGridTools.cpp:397-505, :709-827) so that areas,
//     normals, volumes and centroids are bit-identical.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>

#include "psst_host.hpp"

namespace psst {

namespace {

struct FaceGeom {
    double n[3], A, c[3];
};

// psstFace::initialize (psstFiniteVolumeGridTools.cpp:397-505) followed by
// psstGeometryVector::initialize/normalize (:248-290).
FaceGeom face_geometry(const RawMesh &m, int64_t fi) {
    FaceGeom g{};
    const int dim = m.dim;
    const int nv = m.face_nv[(size_t)fi];
    auto V = [&](int j, int k) { return m.xyz[3 * (size_t)(m.face_v[4 * (size_t)fi + j] - 1) + k]; };
    g.A = 0.0;
    if (dim == 2) {
        if (nv != 2) throw std::runtime_error("Error: Expected to vertices but extra data found!");
        for (int i = 0; i < 2; ++i) {
            g.c[i] = 0.5 * (V(1, i) + V(0, i));
            double comp = V(1, i) - V(0, i);
            g.A += comp * comp;
        }
        g.A = std::sqrt(g.A);
        g.n[0] = V(0, 1) - V(1, 1);
        g.n[1] = V(1, 0) - V(0, 0);
    } else {
        if (nv < 3) throw std::runtime_error("Error: Expected at least 3 vertices!");
        double avg[3] = {0, 0, 0}, avg3[3];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < nv; ++j) avg[i] += V(j, i);
            avg[i] /= (double)nv;
            avg3[i] = avg[i] * 3;
        }
        double a[3], b[3], c[3], ts;
        int last = nv - 1;
        for (int i = 0; i <= last; ++i) {
            int nx = (i == last) ? 0 : i + 1;
            for (int j = 0; j < 3; ++j) {
                a[j] = V(i, j) - avg[j];
                b[j] = V(nx, j) - avg[j];
            }
            c[0] = a[1] * b[2] - a[2] * b[1];
            c[1] = a[2] * b[0] - a[0] * b[2];
            c[2] = a[0] * b[1] - a[1] * b[0];
            ts = 0.0;
            for (int j = 0; j < 3; ++j) ts += c[j] * c[j];
            ts = std::sqrt(ts);
            g.A += ts;
            if (i < last) {
                for (int j = 0; j < 3; ++j) g.c[j] += ts * (a[j] + b[j] + avg3[j]);
            } else {
                double sq3 = g.A * 3;
                for (int j = 0; j < 3; ++j) {
                    g.c[j] += ts * (a[j] + b[j] + avg3[j]);
                    g.c[j] /= sq3;
                    g.n[j] = c[j];  // quirk Q17: normal of the last sub-triangle
                }
            }
        }
        g.A *= 0.5;
    }
    double len = 0.0;
    for (int i = 0; i < dim; ++i) len += g.n[i] * g.n[i];
    len = std::sqrt(len);
    if (std::abs(len) < 1.0e-10) throw std::runtime_error("Error: Division by zero detected!");
    for (int i = 0; i < dim; ++i) g.n[i] /= len;
    return g;
}

}  // namespace

Grid build_grid(const RawMesh &m) {
    Grid g;
    g.dim = m.dim;
    const int dim = m.dim;
    // ---- zones in ascending id (std::map order) ----
    std::map<int64_t, const RawZone *> by_id;
    for (const RawZone &z : m.zones) {
        if (z.first == 0 && z.last == 0) continue;  // name-only zone without data
        by_id[z.id] = &z;
    }
    std::map<int64_t, int> zone_index;
    for (auto &kv : by_id) {
        GridZone gz;
        gz.id = kv.first;
        gz.bc = kv.second->bc;
        gz.name = kv.second->name;
        zone_index[kv.first] = (int)g.zones.size();
        g.zones.push_back(gz);
    }
    // ---- canonical cells: interior zones then ghost zones, ascending id ----
    int32_t n_int = 0, n_gh = 0;
    for (GridZone &z : g.zones) {
        const RawZone &rz = *by_id[z.id];
        if (z.bc == 1) {
            z.has_cells = true;
            z.cell_begin = n_int;
            z.cell_count = (int32_t)(rz.last - rz.first + 1);
            n_int += z.cell_count;
        }
    }
    for (GridZone &z : g.zones) {
        const RawZone &rz = *by_id[z.id];
        if (z.bc != 1 && z.bc != 2 && z.bc != 12 && z.bc != 31 && z.bc != 32) {
            z.has_cells = true;
            z.cell_begin = n_int + n_gh;
            z.cell_count = (int32_t)(rz.last - rz.first + 1);
            n_gh += z.cell_count;
        }
    }
    g.N = n_int;
    g.G = n_gh;
    const int64_t NT = (int64_t)n_int + n_gh;
    int64_t n_parent_cells = 0;
    for (const GridZone &z : g.zones)
        if (z.bc == 32) n_parent_cells += by_id[z.id]->last - by_id[z.id]->first + 1;
    if (n_int + n_parent_cells != m.n_cells()) throw std::runtime_error("mesh: cells outside fluid zones");
    // file cell index (1-based) -> canonical
    std::vector<int32_t> canon_cell((size_t)m.n_cells(), -1);
    for (const GridZone &z : g.zones) {
        if (z.bc != 1) continue;
        const RawZone &rz = *by_id[z.id];
        for (int64_t j = 0; j < z.cell_count; ++j) canon_cell[(size_t)(rz.first - 1 + j)] = z.cell_begin + (int32_t)j;
    }
    // ---- canonical faces: face zones ascending id ----
    int32_t nf = 0;
    int64_t n_parent_faces = 0;
    for (GridZone &z : g.zones) {
        if (z.bc == 1 || z.bc == 32) continue;
        if (z.bc == 31) {
            n_parent_faces += by_id[z.id]->last - by_id[z.id]->first + 1;
            continue;
        }
        const RawZone &rz = *by_id[z.id];
        z.has_faces = true;
        z.face_begin = nf;
        z.face_count = (int32_t)(rz.last - rz.first + 1);
        nf += z.face_count;
    }
    g.F = nf;
    if (nf + n_parent_faces != m.n_faces()) throw std::runtime_error("mesh: faces outside face zones");
    // file face (0-based) -> canonical; -1 for parent faces
    std::vector<int32_t> canon_face((size_t)m.n_faces(), -1);
    std::vector<int32_t> face_local((size_t)nf);  // canonical -> zone-local index
    for (const GridZone &z : g.zones) {
        if (!z.has_faces) continue;
        const RawZone &rz = *by_id[z.id];
        for (int32_t j = 0; j < z.face_count; ++j) {
            canon_face[(size_t)(rz.first - 1 + j)] = z.face_begin + j;
            face_local[(size_t)(z.face_begin + j)] = j;
        }
    }
    // ---- neighbours ----
    g.f0.assign((size_t)nf, -1);
    g.f1.assign((size_t)nf, -1);
    g.ghost_face.assign((size_t)n_gh, -1);
    g.ghost_zone.assign((size_t)n_gh, -1);
    for (const GridZone &z : g.zones) {
        if (!z.has_faces) continue;
        const RawZone &rz = *by_id[z.id];
        int zi = zone_index[z.id];
        for (int32_t j = 0; j < z.face_count; ++j) {
            size_t fi = (size_t)(rz.first - 1 + j);
            int32_t cf = z.face_begin + j;
            for (int k = 0; k < 2; ++k) {
                int64_t cell = m.face_c[2 * fi + k];
                int32_t ci;
                if (cell == 0) {
                    if (z.bc == 2 || z.bc == 12) throw std::runtime_error("mesh: interior face without two cells");
                    ci = z.cell_begin + j;
                    g.ghost_face[(size_t)(ci - n_int)] = cf;
                    g.ghost_zone[(size_t)(ci - n_int)] = zi;
                } else {
                    if (cell < 1 || cell > m.n_cells()) throw std::runtime_error("mesh: face cell index out of range");
                    ci = canon_cell[(size_t)cell - 1];
                    if (ci < 0) throw std::runtime_error("mesh: active face next to a parent cell");
                }
                (k == 0 ? g.f0 : g.f1)[(size_t)cf] = ci;
            }
            if (z.bc != 2 && z.bc != 12) {
                bool g0 = g.f0[(size_t)cf] >= n_int, g1 = g.f1[(size_t)cf] >= n_int;
                if (g0 == g1) throw std::runtime_error("mesh: boundary face must have exactly one ghost");
            }
        }
    }
    for (int32_t gi = 0; gi < n_gh; ++gi)
        if (g.ghost_face[(size_t)gi] < 0) throw std::runtime_error("mesh: ghost without face");
    g.fcell.resize(2 * (size_t)nf);
    std::copy(g.f0.begin(), g.f0.end(), g.fcell.begin());
    std::copy(g.f1.begin(), g.f1.end(), g.fcell.begin() + nf);
    // ---- face-list order: face zones in FILE order, faces in zone order ----
    std::vector<int32_t> cnt((size_t)NT + 1, 0);
    for (int32_t f = 0; f < nf; ++f) {
        cnt[(size_t)g.f0[(size_t)f] + 1]++;
        cnt[(size_t)g.f1[(size_t)f] + 1]++;
    }
    g.list_off.assign((size_t)NT + 1, 0);
    for (int64_t c = 0; c < NT; ++c) g.list_off[(size_t)c + 1] = g.list_off[(size_t)c] + cnt[(size_t)c + 1];
    g.list_face.assign((size_t)g.list_off[(size_t)NT], -1);
    {
        std::vector<int32_t> fill(g.list_off.begin(), g.list_off.end() - 1);
        for (const RawZone &rz : m.zones) {
            if (rz.bc == 1 || rz.bc == 31 || rz.bc == 32 || (rz.first == 0 && rz.last == 0)) continue;
            for (int64_t i = rz.first; i <= rz.last; ++i) {
                int32_t cf = canon_face[(size_t)i - 1];
                g.list_face[(size_t)fill[(size_t)g.f0[(size_t)cf]]++] = cf;
                g.list_face[(size_t)fill[(size_t)g.f1[(size_t)cf]]++] = cf;
            }
        }
    }
    // ---- visit order: canonical (ascending) face index, interior cells ----
    std::vector<int32_t> vc((size_t)n_int + 1, 0);
    for (int32_t f = 0; f < nf; ++f) {
        if (g.f0[(size_t)f] < n_int) vc[(size_t)g.f0[(size_t)f] + 1]++;
        if (g.f1[(size_t)f] < n_int) vc[(size_t)g.f1[(size_t)f] + 1]++;
    }
    g.visit_off.assign((size_t)n_int + 1, 0);
    for (int32_t c = 0; c < n_int; ++c) g.visit_off[(size_t)c + 1] = g.visit_off[(size_t)c] + vc[(size_t)c + 1];
    g.visit_face.assign((size_t)g.visit_off[(size_t)n_int], -1);
    {
        std::vector<int32_t> fill(g.visit_off.begin(), g.visit_off.end() - 1);
        for (int32_t f = 0; f < nf; ++f) {
            if (g.f0[(size_t)f] < n_int) g.visit_face[(size_t)fill[(size_t)g.f0[(size_t)f]]++] = f;
            if (g.f1[(size_t)f] < n_int) g.visit_face[(size_t)fill[(size_t)g.f1[(size_t)f]]++] = f;
        }
    }
    // ---- face geometry ----
    g.fn.assign(3 * (size_t)nf, 0.0);
    g.fc.assign(3 * (size_t)nf, 0.0);
    g.fa.assign((size_t)nf, 0.0);
    for (int64_t fi = 0; fi < m.n_faces(); ++fi) {
        if (canon_face[(size_t)fi] < 0) continue;
        FaceGeom fg = face_geometry(m, fi);
        size_t cf = (size_t)canon_face[(size_t)fi];
        for (int k = 0; k < 3; ++k) {
            g.fn[(size_t)k * nf + cf] = fg.n[k];
            g.fc[(size_t)k * nf + cf] = fg.c[k];
        }
        g.fa[cf] = fg.A;
    }
    // ---- periodic faces (zone type 12, no reference counterpart): the mesh
    //      is periodic along z with period P = its z extent; a periodic
    //      face's vertices lie in cell(0)'s frame and cell(1) sees the face
    //      translated by +P.  fshift = -P carries cell(1) into cell(0)'s frame. ----
    g.fshift.assign(3 * (size_t)nf, 0.0);
    g.periodic = false;
    {
        double zmin = 1e300, zmax = -1e300;
        for (const GridZone &z : g.zones) {
            if (z.bc != 12 || !z.has_faces) continue;
            if (dim != 3) throw std::runtime_error("mesh: periodic zones need a 3D mesh");
            if (!g.periodic) {
                for (int64_t v = 0; v < m.n_vertices(); ++v) {
                    zmin = std::min(zmin, m.xyz[3 * (size_t)v + 2]);
                    zmax = std::max(zmax, m.xyz[3 * (size_t)v + 2]);
                }
                g.periodic = true;
            }
            for (int32_t j = 0; j < z.face_count; ++j) g.fshift[2 * (size_t)nf + (size_t)(z.face_begin + j)] = zmin - zmax;
        }
    }
    // ---- cell geometry (psstVolume::initialize, face-list order) ----
    g.vol.assign((size_t)NT, 0.0);
    g.cx.assign(3 * (size_t)NT, 0.0);
    std::vector<size_t> file_face_of((size_t)nf);
    for (int64_t fi = 0; fi < m.n_faces(); ++fi)
        if (canon_face[(size_t)fi] >= 0) file_face_of[(size_t)canon_face[(size_t)fi]] = (size_t)fi;
    // vertex j of face cf as cell c sees it (a periodic face translated by
    // +P for its cell(1)); FC likewise for the face centre
    auto off = [&](int32_t cf, int64_t c, int k) {
        return (g.periodic && g.f1[(size_t)cf] == c) ? -g.fshift[(size_t)k * nf + cf] : 0.0;
    };
    auto VXc = [&](int32_t cf, int64_t c, int j, int k) {
        size_t fi = file_face_of[(size_t)cf];
        const double x = m.xyz[3 * (size_t)(m.face_v[4 * fi + j] - 1) + k];
        const double o = off(cf, c, k);
        return o != 0.0 ? x + o : x;
    };
    auto FC = [&](int32_t cf, int64_t c, int k) {
        const double x = g.fc[(size_t)k * nf + cf];
        const double o = off(cf, c, k);
        return o != 0.0 ? x + o : x;
    };
    for (int64_t c = 0; c < NT; ++c) {
        auto VX = [&](int32_t cf, int j, int k) { return VXc(cf, c, j, k); };
        int32_t b = g.list_off[(size_t)c], e = g.list_off[(size_t)c + 1];
        if (c >= n_int) {
            int32_t cf = g.ghost_face[(size_t)(c - n_int)];
            for (int k = 0; k < 3; ++k) g.cx[(size_t)k * NT + c] = g.fc[(size_t)k * nf + cf];
            g.vol[(size_t)c] = 0.0;
            continue;
        }
        size_t nfc = (size_t)(e - b);
        double V = 0.0, x[3] = {0, 0, 0};
        if (dim == 2) {
            double avg[2] = {0, 0}, avg3[2];
            for (int j = 0; j < 2; ++j) {
                for (int32_t i = b; i < e; ++i) avg[j] += g.fc[(size_t)j * nf + g.list_face[(size_t)i]];
                avg[j] /= (double)nfc;
                avg3[j] = avg[j] * 3;
            }
            for (int32_t i = b; i < e; ++i) {
                int32_t cf = g.list_face[(size_t)i];
                double a[2], bb[2];
                for (int k = 0; k < 2; ++k) {
                    a[k] = VX(cf, 0, k) - avg[k];
                    bb[k] = VX(cf, 1, k) - avg[k];
                }
                double tv = std::abs(a[0] * bb[1] - a[1] * bb[0]);
                V += tv;
                for (int k = 0; k < 2; ++k) x[k] += tv * (a[k] + bb[k] + avg3[k]);
            }
            double v3 = V * 3;
            for (int k = 0; k < 2; ++k) x[k] /= v3;
            V /= 2;
        } else {
            double avg[3] = {0, 0, 0}, avg4[3];
            for (int j = 0; j < 3; ++j) {
                for (int32_t i = b; i < e; ++i) avg[j] += FC(g.list_face[(size_t)i], c, j);
                avg[j] /= (double)nfc;
                avg4[j] = avg[j] * 4;
            }
            for (int32_t i = b; i < e; ++i) {
                int32_t cf = g.list_face[(size_t)i];
                int nv = m.face_nv[file_face_of[(size_t)cf]];
                for (int j = 0; j < nv; ++j) {
                    int nx = (j + 1 == nv) ? 0 : j + 1;
                    double a[3], bb[3], cc[3];
                    for (int k = 0; k < 3; ++k) {
                        a[k] = VX(cf, j, k) - avg[k];
                        bb[k] = VX(cf, nx, k) - avg[k];
                        cc[k] = FC(cf, c, k) - avg[k];
                    }
                    double tv = std::abs(a[0] * (bb[1] * cc[2] - bb[2] * cc[1]) - a[1] * (bb[0] * cc[2] - bb[2] * cc[0]) +
                                         a[2] * (bb[0] * cc[1] - bb[1] * cc[0]));
                    V += tv;
                    for (int k = 0; k < 3; ++k) x[k] += tv * (a[k] + bb[k] + cc[k] + avg4[k]);
                }
            }
            double v4 = V * 4.0;
            for (int k = 0; k < 3; ++k) x[k] /= v4;
            V /= 6.0;
        }
        g.vol[(size_t)c] = V;
        for (int k = 0; k < 3; ++k) g.cx[(size_t)k * NT + c] = x[k];
    }
    double mn = -1.0;
    for (int32_t c = 0; c < n_int; ++c)
        if (mn > g.vol[(size_t)c] || mn < 0.0) mn = g.vol[(size_t)c];
    g.min_vol = mn;
    // ---- Green-Gauss beta, indexed by zone-local face index (quirk Q1,
    //      psstGradCalculation.cpp:30-70): the last face zone in ascending id
    //      order that has a face with local index i defines beta[i].  Periodic
    //      faces (no reference counterpart) keep their own beta. ----
    {
        int32_t max_local = 0;
        for (const GridZone &z : g.zones)
            if (z.has_faces) max_local = std::max(max_local, z.face_count);
        std::vector<double> beta_local((size_t)max_local, 0.0);
        std::vector<std::pair<int32_t, double>> own;  // periodic faces keep their own beta
        for (const GridZone &z : g.zones) {
            if (!z.has_faces) continue;
            for (int32_t i = 0; i < z.face_count; ++i) {
                int32_t cf = z.face_begin + i;
                double d[2];
                int32_t nb[2] = {g.f0[(size_t)cf], g.f1[(size_t)cf]};
                for (int j = 0; j < 2; ++j) {
                    d[j] = 0.0;
                    if (nb[j] < n_int) {
                        double dist = 0.0;
                        for (int k = 0; k < dim; ++k) {
                            double xc = g.cx[(size_t)k * NT + nb[j]];
                            if (j == 1 && g.periodic) xc = xc + g.fshift[(size_t)k * nf + cf];
                            double add = (g.fc[(size_t)k * nf + cf] - xc);
                            add = add * add;
                            dist += add;
                        }
                        d[j] = std::sqrt(dist);
                    }
                }
                beta_local[(size_t)i] = d[1] / (d[0] + d[1]);
                if (z.bc == 12) own.push_back({cf, beta_local[(size_t)i]});
            }
        }
        g.beta.assign((size_t)nf, 0.0);
        for (int32_t f = 0; f < nf; ++f) g.beta[(size_t)f] = beta_local[(size_t)face_local[(size_t)f]];
        for (auto &fb : own) g.beta[(size_t)fb.first] = fb.second;
    }
    // ---- least-squares factors, indexed by zone-local cell index (quirk Q20,
    //      psstGradCalculation.cpp:191-260) ----
    {
        int32_t max_local = 0;
        for (const GridZone &z : g.zones)
            if (z.bc == 1) max_local = std::max(max_local, z.cell_count);
        std::vector<double> r_local(6 * (size_t)max_local, 0.0), b_local((size_t)max_local, 0.0);
        g.cell_zone_local.assign((size_t)n_int, 0);
        for (const GridZone &z : g.zones) {
            if (z.bc != 1) continue;
            for (int32_t i = 0; i < z.cell_count; ++i) {
                int32_t c = z.cell_begin + i;
                g.cell_zone_local[(size_t)c] = i;
                double cf[6] = {0, 0, 0, 0, 0, 0};
                for (int32_t l = g.list_off[(size_t)c]; l < g.list_off[(size_t)c + 1]; ++l) {
                    int32_t f = g.list_face[(size_t)l];
                    const bool c_is_0 = g.f0[(size_t)f] == c;
                    int32_t nb = c_is_0 ? g.f1[(size_t)f] : g.f0[(size_t)f];
                    double diff[3] = {0, 0, 0}, w = 0.0;
                    for (int k = 0; k < dim; ++k) {
                        double xn = g.cx[(size_t)k * NT + nb];
                        if (g.periodic) {
                            const double sh = g.fshift[(size_t)k * nf + f];
                            if (sh != 0.0) xn = c_is_0 ? xn + sh : xn - sh;
                        }
                        diff[k] = xn - g.cx[(size_t)k * NT + c];
                        w += diff[k] * diff[k];
                    }
                    double ws = 1.0 / w;
                    cf[0] += ws * diff[0] * diff[0];
                    cf[1] += ws * diff[0] * diff[1];
                    cf[2] += ws * diff[1] * diff[1];
                    if (dim == 3) {
                        cf[3] += ws * diff[0] * diff[2];
                        cf[4] += ws * diff[1] * diff[2];
                        cf[5] += ws * diff[2] * diff[2];
                    }
                }
                double *r = &r_local[6 * (size_t)i];
                r[0] = std::sqrt(cf[0]);
                r[1] = cf[1] / r[0];
                r[2] = std::sqrt(cf[2] - r[1] * r[1]);
                b_local[(size_t)i] = 0.0;
                if (dim == 3) {
                    r[3] = cf[3] / r[0];
                    r[4] = (cf[4] - r[1] / r[0] * cf[3]) / r[2];
                    r[5] = std::sqrt(cf[5] - (r[3] * r[3] + r[4] * r[4]));
                    b_local[(size_t)i] = (r[1] * r[4] - r[3] * r[2]) / (r[0] * r[2]);
                }
            }
        }
        g.lsq_r.assign(6 * (size_t)n_int, 0.0);
        g.lsq_beta.assign((size_t)n_int, 0.0);
        for (int32_t c = 0; c < n_int; ++c) {
            int32_t i = g.cell_zone_local[(size_t)c];
            for (int k = 0; k < 6; ++k) g.lsq_r[(size_t)k * n_int + c] = r_local[6 * (size_t)i + k];
            g.lsq_beta[(size_t)c] = b_local[(size_t)i];
        }
    }
    return g;
}

psst_mesh_desc make_mesh_desc(const Grid &g) {
    // The device takes [dim][n] rows; the grid stores 3 rows, of which the
    // first dim are used (rows are n apart).
    psst_mesh_desc d{};
    d.dim = g.dim;
    d.n_cells = g.N;
    d.n_ghosts = g.G;
    d.n_faces = g.F;
    d.cell_volume = g.vol.data();
    d.cell_centre = g.cx.data();
    d.face_normal = g.fn.data();
    d.face_area = g.fa.data();
    d.face_cells = g.fcell.data();
    d.face_beta = g.beta.data();
    d.visit_offsets = g.visit_off.data();
    d.visit_faces = g.visit_face.data();
    d.list_offsets = g.list_off.data();
    d.list_faces = g.list_face.data();
    d.ghost_face = g.ghost_face.data();
    d.ghost_bc = g.ghost_zone.data();
    d.lsq_r = g.lsq_r.data();
    d.lsq_beta = g.lsq_beta.data();
    d.min_volume = g.min_vol;
    d.face_shift = g.periodic ? g.fshift.data() : nullptr;
    return d;
}

}  // namespace psst
