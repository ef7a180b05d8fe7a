/*
 * psst_ref_adapter.hpp -- header-only adapter for the reference's own build
 * (PSST, nkmashaev/SupersonicFlowSolver).  It keeps the reference's Fluent
 * reader and its finite-volume grid: the grid that
 * psstAnsysInputGrid::create_finite_volume_grid() returns (main.cpp:46-50) is
 * flattened from its storages, psstGridFV::face_storage() and
 * volume_storage() (psstFiniteVolumeGrid.h:59-65, psstFaceStorage /
 * psstVolumeStorage psstFiniteVolumeGridTools.h:336-360), into the
 * psst_mesh_desc of include/psst_gpu.h, with the reference's own geometry
 * (psstFace::initialize / psstVolume::initialize, computed by the reference).
 * It also moves the primitive state between the reference's psstVariable
 * objects and the [(dim+3)][N+G] arrays of the C-ABI, so the reference's
 * save_data() writes the device's solution.
 *
 * Include it after the reference's headers, in a translation unit of the
 * reference's build.  Orders produced (those include/psst_gpu.h documents):
 *   cells: volume zones of Fluent type 1 in ascending zone id (the map order
 *          of volume_storage), each in its storage order; then ghost cells,
 *          boundary zones in ascending id, one per boundary face in face order;
 *   faces: face zones in ascending id (face_storage's map order), parent
 *          zones (type 31) skipped, each in its storage order;
 *   visit order of a cell's faces: ascending canonical face (the flux and
 *          Green-Gauss loops, psstConvectiveFlow.cpp:118-123);
 *   list order: the cell's own face list, psstVolume::face(i) (the time step
 *          and least-squares sums).
 * The Green-Gauss weights and least-squares factors follow
 * psstGradCalculation.cpp:17-72 and :191-260, including their zone-local
 * indexing (a weight or factor of local index i is the one of the last zone
 * in ascending id order that has an element i).
 */
#ifndef PSST_REF_ADAPTER_HPP
#define PSST_REF_ADAPTER_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "psst_gpu.h"

struct PsstRefMesh {
    int32_t dim = 2, N = 0, G = 0, F = 0;
    std::vector<double> vol, cx, fn, fa, beta, lsq_r, lsq_beta;  // rows of n values, as psst_mesh_desc
    std::vector<int32_t> fcell, voff, vface, loff, lface, gface, gzone;
    double min_volume = 0.0;
    // zone table in ascending id (psst_case_from_mesh); names kept alive here
    std::vector<std::string> zone_names;
    std::vector<psst_zone> zones;
    // canonical cell -> the reference's volume (interior cells, then ghosts)
    std::vector<std::shared_ptr<psstInterfaceVolume>> cells;

    psst_mesh_desc desc() const {
        psst_mesh_desc d{};
        d.dim = dim;
        d.n_cells = N;
        d.n_ghosts = G;
        d.n_faces = F;
        d.cell_volume = vol.data();
        d.cell_centre = cx.data();
        d.face_normal = fn.data();
        d.face_area = fa.data();
        d.face_cells = fcell.data();
        d.face_beta = beta.data();
        d.visit_offsets = voff.data();
        d.visit_faces = vface.data();
        d.list_offsets = loff.data();
        d.list_faces = lface.data();
        d.ghost_face = gface.data();
        d.ghost_bc = gzone.data();
        d.lsq_r = lsq_r.data();
        d.lsq_beta = lsq_beta.data();
        d.min_volume = min_volume;
        d.face_shift = nullptr;  // the reference has no periodic faces
        return d;
    }
};

inline PsstRefMesh psst_ref_flatten(psstGridFV &grid) {
    PsstRefMesh m;
    const int dim = (int)grid.dimension();
    m.dim = dim;
    std::map<size_t, psstFaceStorage> &fs = *grid.face_storage();
    std::map<size_t, psstVolumeStorage> &vs = *grid.volume_storage();
    // ---- zones: every id of either map, ascending ----
    std::map<size_t, psst_zone> zmap;
    std::map<size_t, std::string> zname;
    for (auto &kv : vs) {
        psst_zone &z = zmap[kv.first];
        z.id = (int64_t)kv.first;
        z.type = (int32_t)kv.second.bc_type_id_;
        if (z.type == 1 || z.type == 32) z.count = (int32_t)kv.second.volumes_.size();
        zname[kv.first] = kv.second.name_;
    }
    for (auto &kv : fs) {
        psst_zone &z = zmap[kv.first];
        z.id = (int64_t)kv.first;
        z.type = (int32_t)kv.second.bc_type_id_;
        z.count = (int32_t)kv.second.faces_.size();
        zname[kv.first] = kv.second.name_;
    }
    std::map<size_t, int32_t> zindex;
    for (auto &kv : zmap) {
        if (kv.second.count == 0) continue;  // a zone without elements
        zindex[kv.first] = (int32_t)m.zones.size();
        m.zone_names.push_back(zname[kv.first]);
        m.zones.push_back(kv.second);
    }
    for (size_t i = 0; i < m.zones.size(); ++i) m.zones[i].name = m.zone_names[i].c_str();
    auto is_ghost_zone = [](int32_t t) { return t != 1 && t != 2 && t != 12 && t != 31 && t != 32; };
    // ---- canonical cells ----
    std::unordered_map<const psstInterfaceVolume *, int32_t> cell_of;
    for (auto &kv : vs)
        if (kv.second.bc_type_id_ == 1)
            for (auto &v : kv.second.volumes_) {
                cell_of[v.get()] = (int32_t)m.cells.size();
                m.cells.push_back(v);
            }
    m.N = (int32_t)m.cells.size();
    std::vector<int32_t> ghost_zone_of;
    for (auto &kv : vs)
        if (is_ghost_zone((int32_t)kv.second.bc_type_id_))
            for (auto &v : kv.second.volumes_) {
                cell_of[v.get()] = (int32_t)m.cells.size();
                m.cells.push_back(v);
                ghost_zone_of.push_back(zindex.at(kv.first));
            }
    m.G = (int32_t)m.cells.size() - m.N;
    const int32_t NT = m.N + m.G;
    // ---- canonical faces ----
    std::vector<std::shared_ptr<psstInterfaceFace>> faces;
    std::vector<int32_t> face_local;  // zone-local index (Green-Gauss weights)
    std::unordered_map<const psstInterfaceFace *, int32_t> face_of;
    for (auto &kv : fs) {
        if (kv.second.bc_type_id_ == 31) continue;
        int32_t j = 0;
        for (auto &f : kv.second.faces_) {
            face_of[f.get()] = (int32_t)faces.size();
            faces.push_back(f);
            face_local.push_back(j++);
        }
    }
    m.F = (int32_t)faces.size();
    const int32_t F = m.F;
    // ---- neighbours and face geometry ----
    m.fcell.assign(2 * (size_t)F, -1);
    m.fn.assign((size_t)dim * F, 0.0);
    m.fa.assign((size_t)F, 0.0);
    std::vector<double> fc((size_t)dim * F, 0.0);
    for (int32_t f = 0; f < F; ++f) {
        const psstInterfaceFace &face = *faces[(size_t)f];
        for (int k = 0; k < 2; ++k) {
            auto it = cell_of.find(face.neighbour_ptr((size_t)k).get());
            if (it == cell_of.end()) throw std::runtime_error("psst_ref_flatten: a face next to no solved cell");
            m.fcell[(size_t)k * F + f] = it->second;
        }
        for (int k = 0; k < dim; ++k) {
            m.fn[(size_t)k * F + f] = face.normal_vector()->component((size_t)k);
            fc[(size_t)k * F + f] = face.center_vertex()->component((size_t)k);
        }
        m.fa[(size_t)f] = face.square();
    }
    // ---- cell geometry ----
    m.vol.assign((size_t)NT, 0.0);
    m.cx.assign((size_t)dim * NT, 0.0);
    for (int32_t c = 0; c < NT; ++c) {
        const psstInterfaceVolume &v = *m.cells[(size_t)c];
        m.vol[(size_t)c] = v.volume();
        for (int k = 0; k < dim; ++k) m.cx[(size_t)k * NT + c] = v.center_vertex()->component((size_t)k);
    }
    m.min_volume = grid.min_vol();
    // ---- per-cell face lists: visit order and the cell's own list ----
    m.voff.assign((size_t)m.N + 1, 0);
    for (int32_t f = 0; f < F; ++f)
        for (int k = 0; k < 2; ++k)
            if (m.fcell[(size_t)k * F + f] < m.N) m.voff[(size_t)m.fcell[(size_t)k * F + f] + 1]++;
    for (int32_t c = 0; c < m.N; ++c) m.voff[(size_t)c + 1] += m.voff[(size_t)c];
    m.vface.assign((size_t)m.voff[(size_t)m.N], -1);
    {
        std::vector<int32_t> fill(m.voff.begin(), m.voff.end() - 1);
        for (int32_t f = 0; f < F; ++f)
            for (int k = 0; k < 2; ++k) {
                const int32_t c = m.fcell[(size_t)k * F + f];
                if (c < m.N) m.vface[(size_t)fill[(size_t)c]++] = f;
            }
    }
    m.loff.assign((size_t)m.N + 1, 0);
    for (int32_t c = 0; c < m.N; ++c) {
        const psstInterfaceVolume &v = *m.cells[(size_t)c];
        m.loff[(size_t)c + 1] = m.loff[(size_t)c] + (int32_t)v.number_of_faces();
        for (size_t i = 0; i < v.number_of_faces(); ++i) m.lface.push_back(face_of.at(v.face_ptr(i).get()));
    }
    m.gface.assign((size_t)m.G, -1);
    m.gzone = ghost_zone_of;
    for (int32_t g = 0; g < m.G; ++g) {
        const psstInterfaceVolume &v = *m.cells[(size_t)(m.N + g)];
        if (v.number_of_faces() != 1) throw std::runtime_error("psst_ref_flatten: a ghost cell with other than one face");
        m.gface[(size_t)g] = face_of.at(v.face_ptr(0).get());
    }
    // ---- Green-Gauss weights beta = d1 / (d0 + d1) (0 for a ghost), by
    //      zone-local face index (psstGradCalculation.cpp:30-70) ----
    {
        const int32_t nloc = face_local.empty() ? 0 : *std::max_element(face_local.begin(), face_local.end()) + 1;
        std::vector<double> local((size_t)nloc, 0.0);
        for (int32_t f = 0; f < F; ++f) {
            double d[2];
            for (int j = 0; j < 2; ++j) {
                const int32_t c = m.fcell[(size_t)j * F + f];
                d[j] = 0.0;
                if (c < m.N) {
                    double dist = 0.0;
                    for (int k = 0; k < dim; ++k) {
                        double add = fc[(size_t)k * F + f] - m.cx[(size_t)k * NT + c];
                        add = add * add;
                        dist += add;
                    }
                    d[j] = std::sqrt(dist);
                }
            }
            local[(size_t)face_local[(size_t)f]] = d[1] / (d[0] + d[1]);
        }
        m.beta.assign((size_t)F, 0.0);
        for (int32_t f = 0; f < F; ++f) m.beta[(size_t)f] = local[(size_t)face_local[(size_t)f]];
    }
    // ---- least-squares factors by zone-local cell index
    //      (psstGradCalculation.cpp:191-260) ----
    {
        std::vector<int32_t> cell_local((size_t)m.N, 0);
        int32_t c0 = 0, nloc = 0;
        for (auto &kv : vs)
            if (kv.second.bc_type_id_ == 1) {
                const int32_t n = (int32_t)kv.second.volumes_.size();
                for (int32_t i = 0; i < n; ++i) cell_local[(size_t)(c0 + i)] = i;
                c0 += n;
                nloc = std::max(nloc, n);
            }
        std::vector<double> r_local(6 * (size_t)nloc, 0.0), b_local((size_t)nloc, 0.0);
        for (int32_t c = 0; c < m.N; ++c) {
            double cf[6] = {0, 0, 0, 0, 0, 0};
            for (int32_t l = m.loff[(size_t)c]; l < m.loff[(size_t)c + 1]; ++l) {
                const int32_t f = m.lface[(size_t)l];
                const int32_t nb = m.fcell[(size_t)f] == c ? m.fcell[(size_t)F + f] : m.fcell[(size_t)f];
                double diff[3] = {0, 0, 0}, w = 0.0;
                for (int k = 0; k < dim; ++k) {
                    diff[k] = m.cx[(size_t)k * NT + nb] - m.cx[(size_t)k * NT + c];
                    w += diff[k] * diff[k];
                }
                const double ws = 1.0 / w;
                cf[0] += ws * diff[0] * diff[0];
                cf[1] += ws * diff[0] * diff[1];
                cf[2] += ws * diff[1] * diff[1];
                if (dim == 3) {
                    cf[3] += ws * diff[0] * diff[2];
                    cf[4] += ws * diff[1] * diff[2];
                    cf[5] += ws * diff[2] * diff[2];
                }
            }
            const int32_t i = cell_local[(size_t)c];
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
        m.lsq_r.assign(6 * (size_t)m.N, 0.0);
        m.lsq_beta.assign((size_t)m.N, 0.0);
        for (int32_t c = 0; c < m.N; ++c) {
            const int32_t i = cell_local[(size_t)c];
            for (int k = 0; k < 6; ++k) m.lsq_r[(size_t)k * m.N + c] = r_local[6 * (size_t)i + k];
            m.lsq_beta[(size_t)c] = b_local[(size_t)i];
        }
    }
    return m;
}

// The reference's variables (after solver_parameters() set the initial
// field and the ghost states) as the C-ABI's [(dim+3)][N+G] primitive state:
// u_0..u_{dim-1}, p, T, rho per cell.
inline void psst_ref_read_state(const PsstRefMesh &m, double *prim) {
    const size_t NT = (size_t)m.N + m.G;
    for (size_t c = 0; c < NT; ++c) {
        const psstInterfaceVariable &v = *m.cells[c]->variable();
        for (int k = 0; k < m.dim; ++k) prim[(size_t)k * NT + c] = v.velocity((size_t)k);
        prim[(size_t)m.dim * NT + c] = v.pressure();
        prim[(size_t)(m.dim + 1) * NT + c] = v.temperature();
        prim[(size_t)(m.dim + 2) * NT + c] = v.density();
    }
}

// The device's primitive state back into the reference's variables, so that
// psstGridFV::save_data() writes it (psstFiniteVolumeGrid.cpp:240-246).
inline void psst_ref_write_state(const PsstRefMesh &m, const double *prim) {
    const size_t NT = (size_t)m.N + m.G;
    for (size_t c = 0; c < NT; ++c) {
        psstInterfaceVariable &v = *m.cells[c]->variable();
        for (int k = 0; k < m.dim; ++k) v.velocity((size_t)k) = prim[(size_t)k * NT + c];
        v.pressure() = prim[(size_t)m.dim * NT + c];
        v.temperature() = prim[(size_t)(m.dim + 1) * NT + c];
        v.density() = prim[(size_t)(m.dim + 2) * NT + c];
    }
}

#endif /* PSST_REF_ADAPTER_HPP */
