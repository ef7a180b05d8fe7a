// host_api.cpp -- the host half of the C-ABI (psst_case_*) and the
// psstGridFV driver that replaces the reference's explicit_solve loop with
// one device call per save interval.
//
// psstGridFV mirrors the reference class's entry points
// (psstFiniteVolumeGrid.h:52-85): settings_file_name(), solver_parameters(),
// explicit_solve(), save_data().  Failures surface as std::runtime_error, as
// the reference's own exceptions do.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>

#include "psst_host.hpp"

struct psst_case {
    psst::RawMesh raw;
    psst::Case c;
};

namespace {
thread_local std::string g_case_error;

psst_status fail(const std::exception &e) {
    g_case_error = e.what();
    return PSST_BAD_ARG;
}

psst::WedgeParams to_params(const psst_wedge_params *p) {
    psst::WedgeParams w;
    w.nx = p->nx;
    w.ny = p->ny;
    w.nz = p->nz;
    w.length = p->length;
    w.height = p->height;
    w.depth = p->depth;
    w.ramp_start = p->ramp_start;
    w.ramp_length = p->ramp_length;
    w.ramp_deg = p->ramp_deg;
    w.jitter = p->jitter;
    w.jitter_seed = p->jitter_seed;
    w.numbering = p->numbering;
    w.perm_seed = p->perm_seed;
    w.tets = p->tets;
    w.periodic_z = p->periodic_z;
    return w;
}
}  // namespace

namespace psst {
RawMesh generate_wedge_any(const WedgeParams &p);

void log_phase(const char *what) {
    static const bool on = std::getenv("PSST_VERBOSE") != nullptr;
    if (!on) return;
    static auto last = std::chrono::steady_clock::now();
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "psst: %s (%.1f s)\n", what, std::chrono::duration<double>(now - last).count());
    std::fflush(stderr);
    last = now;
}
}

// ---------------------------------------------------------------------------
// psstGridFV: host driver with the reference's surface.
// ---------------------------------------------------------------------------
class psstGridFV {
   public:
    psstGridFV(psst::Case &c, const psst::RawMesh &raw) : case_(c), raw_(raw) {}
    ~psstGridFV() {
        if (ctx_) psst_ctx_destroy(ctx_);
    }
    void settings_file_name(const std::string &name) { case_.prefix = name; }
    // Creates the device context for the loaded case (the strategy objects
    // of the reference's solver_parameters, psstFiniteVolumeGrid.cpp:36-194).
    void solver_parameters() {
        psst_mesh_desc d = psst::make_mesh_desc(case_.grid);
        case_.settings.bc = case_.bc.data();
        case_.settings.n_bc = (int32_t)case_.bc.size();
        psst_status st = psst_ctx_create(&d, &case_.settings, &ctx_);
        if (st != PSST_OK) throw std::runtime_error(psst_last_error(nullptr));
        check(psst_set_state(ctx_, case_.prim.data()));
    }
    // psstFiniteVolumeGrid.cpp:248-401: iterations iter_..max_iter_, one
    // residual line each, a save every iter_step iterations and at the end.
    void explicit_solve() {
        const int ws = case_.grid.dim + 2;
        std::ofstream out;
        std::string rname = case_.prefix + "_residual.dat";
        if (case_.iter == 1) out.open(rname);
        else out.open(rname, std::ios::app);
        if (!out.is_open()) throw std::runtime_error("Error: Could not open " + rname);
        auto begin = std::chrono::steady_clock::now();
        std::vector<double> res;
        while (case_.iter <= case_.max_iter) {
            // run up to the next save point in one device call
            int64_t next_save = ((case_.iter + case_.iter_step - 1) / case_.iter_step) * case_.iter_step;
            if (next_save < case_.iter) next_save = case_.iter;
            int64_t last = std::min<int64_t>(next_save, case_.max_iter);
            int32_t n = (int32_t)(last - case_.iter + 1);
            res.assign((size_t)n * ws, 0.0);
            int32_t done = 0;
            check(psst_run(ctx_, n, 0.0, res.data(), &done));
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
            for (int32_t k = 0; k < n; ++k) {
                int64_t it = case_.iter + k;
                // the reference truncates elapsed time to milliseconds (:371-373)
                out << psst::residual_line(it, (double)(int64_t)(secs * 1000.0) / 1000.0, &res[(size_t)k * ws], ws);
                if (it % case_.iter_step == 0) {
                    out << std::endl;
                } else {
                    out << "\n";
                }
            }
            case_.iter += n;
            if ((case_.iter - 1) % case_.iter_step == 0) save_data(case_.iter - 1);
        }
        out.close();
        save_data(case_.iter);
    }
    // psstGridFV::save_data (psstFiniteVolumeGrid.cpp:240-246): the output
    // manager's field and case files, then the checkpoint
    void save_data(int64_t iter_value) {
        std::vector<double> prim(case_.prim.size());
        check(psst_get_state(ctx_, prim.data()));
        psst::write_dat(case_.grid, prim.data(), case_.prefix + ".dat");
        psst::write_cas(raw_, case_.prefix + ".cas");
        psst::write_sav(case_, prim.data(), iter_value, case_.prefix + ".sav");
    }

   private:
    void check(psst_status st) {
        if (st != PSST_OK) throw std::runtime_error(psst_last_error(ctx_));
    }
    psst::Case &case_;
    const psst::RawMesh &raw_;
    psst_ctx *ctx_ = nullptr;
};

extern "C" {

const char *psst_case_last_error(void) { return g_case_error.c_str(); }

psst_status psst_case_open(const char *mesh_path, const char *case_prefix, psst_case **out) {
    if (!mesh_path || !case_prefix || !out) return PSST_BAD_ARG;
    try {
        auto pc = std::make_unique<psst_case>();
        pc->raw = psst::read_fluent_msh(mesh_path);
        pc->c.grid = psst::build_grid(pc->raw);
        psst::load_settings(pc->c, case_prefix);
        *out = pc.release();
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

psst_status psst_wedge_write_msh(const psst_wedge_params *p, const char *path) {
    if (!p || !path) return PSST_BAD_ARG;
    try {
        psst::RawMesh m = psst::generate_wedge_any(to_params(p));
        psst::write_fluent_msh(m, path);
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

psst_status psst_case_open_wedge(const psst_wedge_params *p, const char *case_prefix, psst_case **out) {
    if (!p || !case_prefix || !out) return PSST_BAD_ARG;
    try {
        auto pc = std::make_unique<psst_case>();
        psst::log_phase("wedge mesh: generating");
        pc->raw = psst::generate_wedge_any(to_params(p));
        psst::log_phase("wedge mesh: generated; building the grid");
        pc->c.grid = psst::build_grid(pc->raw);
        psst::log_phase("grid built");
        psst::load_settings(pc->c, case_prefix);
        *out = pc.release();
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

psst_status psst_case_from_mesh(const psst_mesh_desc *m, int32_t n_zones, const psst_zone *zones,
                                const char *case_prefix, psst_case **out) {
    if (!m || !zones || n_zones <= 0 || !case_prefix || !out) return PSST_BAD_ARG;
    try {
        auto pc = std::make_unique<psst_case>();
        psst::Grid &g = pc->c.grid;
        const int dim = m->dim;
        if (dim != 2 && dim != 3) throw std::invalid_argument("psst_case_from_mesh: dim must be 2 or 3");
        const int32_t N = m->n_cells, G = m->n_ghosts, F = m->n_faces, NT = N + G;
        if (N <= 0 || G < 0 || F <= 0) throw std::invalid_argument("psst_case_from_mesh: empty mesh");
        g.dim = dim;
        g.N = N;
        g.G = G;
        g.F = F;
        g.vol.assign(m->cell_volume, m->cell_volume + NT);
        g.cx.assign(3 * (size_t)NT, 0.0);
        g.fn.assign(3 * (size_t)F, 0.0);
        for (int k = 0; k < dim; ++k) {
            std::copy(m->cell_centre + (size_t)k * NT, m->cell_centre + (size_t)(k + 1) * NT, g.cx.begin() + (size_t)k * NT);
            std::copy(m->face_normal + (size_t)k * F, m->face_normal + (size_t)(k + 1) * F, g.fn.begin() + (size_t)k * F);
        }
        g.fa.assign(m->face_area, m->face_area + F);
        g.fc.assign(3 * (size_t)F, 0.0);  // face centres are not needed past the geometry the caller computed
        g.fcell.assign(m->face_cells, m->face_cells + 2 * (size_t)F);
        g.f0.assign(m->face_cells, m->face_cells + F);
        g.f1.assign(m->face_cells + F, m->face_cells + 2 * (size_t)F);
        g.beta.assign(m->face_beta, m->face_beta + F);
        g.visit_off.assign(m->visit_offsets, m->visit_offsets + N + 1);
        g.visit_face.assign(m->visit_faces, m->visit_faces + m->visit_offsets[N]);
        g.list_off.assign(m->list_offsets, m->list_offsets + N + 1);
        g.list_face.assign(m->list_faces, m->list_faces + m->list_offsets[N]);
        g.ghost_face.assign(m->ghost_face, m->ghost_face + G);
        g.ghost_zone.assign(m->ghost_bc, m->ghost_bc + G);
        if (m->lsq_r) g.lsq_r.assign(m->lsq_r, m->lsq_r + 6 * (size_t)N);
        if (m->lsq_beta) g.lsq_beta.assign(m->lsq_beta, m->lsq_beta + N);
        g.min_vol = m->min_volume;
        g.periodic = m->face_shift != nullptr;
        if (g.periodic) g.fshift.assign(m->face_shift, m->face_shift + 3 * (size_t)F);
        // zones in ascending id with the canonical ranges their counts give
        int32_t nc = 0, ng = 0, nf = 0;
        for (int32_t i = 0; i < n_zones; ++i) {
            const psst_zone &z = zones[i];
            if (i > 0 && z.id <= zones[i - 1].id) throw std::invalid_argument("psst_case_from_mesh: zones not in ascending id");
            psst::GridZone gz;
            gz.id = z.id;
            gz.bc = z.type;
            gz.name = z.name ? z.name : "";
            if (z.type == 1) {
                gz.has_cells = true;
                gz.cell_begin = nc;
                gz.cell_count = z.count;
                nc += z.count;
            }
            g.zones.push_back(gz);
        }
        for (psst::GridZone &z : g.zones) {
            const int32_t cnt = zones[&z - g.zones.data()].count;
            if (z.bc != 1 && z.bc != 2 && z.bc != 12 && z.bc != 31 && z.bc != 32) {
                z.has_cells = true;
                z.cell_begin = N + ng;
                z.cell_count = cnt;
                ng += cnt;
            }
            if (z.bc != 1 && z.bc != 31 && z.bc != 32) {
                z.has_faces = true;
                z.face_begin = nf;
                z.face_count = cnt;
                nf += cnt;
            }
        }
        if (nc != N || ng != G || nf != F)
            throw std::invalid_argument("psst_case_from_mesh: zone counts do not add up to the mesh's cells, ghosts and faces");
        for (int32_t i = 0; i < G; ++i)
            if (g.ghost_zone[(size_t)i] < 0 || g.ghost_zone[(size_t)i] >= n_zones)
                throw std::invalid_argument("psst_case_from_mesh: ghost_bc out of range");
        psst::load_settings(pc->c, case_prefix);
        *out = pc.release();
        return PSST_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

psst_status psst_case_mesh(const psst_case *c, psst_mesh_desc *mesh) {
    if (!c || !mesh) return PSST_BAD_ARG;
    *mesh = psst::make_mesh_desc(c->c.grid);
    return PSST_OK;
}

psst_status psst_case_settings(const psst_case *c, psst_settings *s) {
    if (!c || !s) return PSST_BAD_ARG;
    *s = c->c.settings;
    s->bc = c->c.bc.data();
    s->n_bc = (int32_t)c->c.bc.size();
    return PSST_OK;
}

psst_status psst_case_set_settings(psst_case *c, const psst_settings *s) {
    if (!c || !s) return PSST_BAD_ARG;
    // a boundary table, when given, replaces the case's entry for entry
    if (s->bc && s->bc != c->c.bc.data() && s->n_bc != (int32_t)c->c.bc.size()) return PSST_BAD_ARG;
    c->c.settings = *s;
    if (s->bc && s->bc != c->c.bc.data())
        std::memcpy(c->c.bc.data(), s->bc, sizeof(psst_bc) * c->c.bc.size());
    c->c.settings.bc = c->c.bc.data();
    c->c.settings.n_bc = (int32_t)c->c.bc.size();
    return PSST_OK;
}

psst_status psst_case_initial_state(const psst_case *c, double *prim) {
    if (!c || !prim) return PSST_BAD_ARG;
    std::memcpy(prim, c->c.prim.data(), sizeof(double) * c->c.prim.size());
    return PSST_OK;
}

psst_status psst_case_info(const psst_case *c, int64_t *info) {
    if (!c || !info) return PSST_BAD_ARG;
    const psst::Grid &g = c->c.grid;
    info[0] = g.dim;
    info[1] = g.N;
    info[2] = g.G;
    info[3] = g.F;
    info[4] = c->c.settings.rk_levels;
    info[5] = c->c.iter;
    info[6] = c->c.max_iter;
    info[7] = c->c.iter_step;
    return PSST_OK;
}

psst_status psst_case_raw_counts(const psst_case *c, int64_t *counts) {
    if (!c || !counts) return PSST_BAD_ARG;
    counts[0] = c->raw.n_vertices();
    counts[1] = c->raw.n_faces();
    counts[2] = c->raw.n_cells();
    counts[3] = (int64_t)c->raw.zones.size();
    return PSST_OK;
}

psst_status psst_case_raw(const psst_case *c, double *xyz, int8_t *face_nv, int64_t *face_v, int64_t *face_c,
                          int64_t *face_zone, int64_t *cell_zone, int64_t *zone_id, int64_t *zone_range,
                          int64_t *zone_bc, char *zone_names, int32_t names_len) {
    if (!c) return PSST_BAD_ARG;
    try {
        const psst::RawMesh &m = c->raw;
        if (xyz) std::memcpy(xyz, m.xyz.data(), sizeof(double) * m.xyz.size());
        if (face_nv) std::memcpy(face_nv, m.face_nv.data(), m.face_nv.size());
        if (face_v) std::memcpy(face_v, m.face_v.data(), sizeof(int64_t) * m.face_v.size());
        if (face_c) std::memcpy(face_c, m.face_c.data(), sizeof(int64_t) * m.face_c.size());
        if (face_zone) std::memcpy(face_zone, m.face_zone.data(), sizeof(int64_t) * m.face_zone.size());
        if (cell_zone) std::memcpy(cell_zone, m.cell_zone.data(), sizeof(int64_t) * m.cell_zone.size());
        std::string names;
        for (size_t i = 0; i < m.zones.size(); ++i) {
            if (zone_id) zone_id[i] = m.zones[i].id;
            if (zone_range) {
                zone_range[2 * i] = m.zones[i].first;
                zone_range[2 * i + 1] = m.zones[i].last;
            }
            if (zone_bc) zone_bc[i] = m.zones[i].bc;
            names += m.zones[i].name;
            names += '\n';
        }
        if (zone_names) {
            if ((int32_t)names.size() + 1 > names_len) return PSST_BAD_ARG;
            std::memcpy(zone_names, names.c_str(), names.size() + 1);
        }
        return PSST_OK;
    } catch (const std::exception &) {  // no exception crosses the C ABI
        return PSST_BAD_ARG;
    }
}

psst_status psst_case_write_sav(const psst_case *c, const double *prim, int64_t iter_value, const char *path) {
    if (!c || !prim || !path) return PSST_BAD_ARG;
    try {
        psst::write_sav(c->c, prim, iter_value, path);
        return PSST_OK;
    } catch (const std::exception &e) {
        g_case_error = e.what();
        return PSST_IO_ERROR;
    }
}

psst_status psst_case_write_dat(const psst_case *c, const double *prim, const char *path) {
    if (!c || !prim || !path) return PSST_BAD_ARG;
    try {
        psst::write_dat(c->c.grid, prim, path);
        return PSST_OK;
    } catch (const std::exception &e) {
        g_case_error = e.what();
        return PSST_IO_ERROR;
    }
}

psst_status psst_case_write_cas(const psst_case *c, const char *path) {
    if (!c || !path) return PSST_BAD_ARG;
    if (c->raw.n_vertices() == 0) {
        g_case_error = "psst_case_write_cas: the case has no Fluent mesh behind it (psst_case_from_mesh)";
        return PSST_UNSUPPORTED;
    }
    try {
        psst::write_cas(c->raw, path);
        return PSST_OK;
    } catch (const std::exception &e) {
        g_case_error = e.what();
        return PSST_IO_ERROR;
    }
}

psst_status psst_case_solve(psst_case *c) {
    if (!c) return PSST_BAD_ARG;
    if (c->raw.n_vertices() == 0) {
        g_case_error = "psst_case_solve: the case has no Fluent mesh behind it (psst_case_from_mesh)";
        return PSST_UNSUPPORTED;
    }
    try {
        psstGridFV fv(c->c, c->raw);
        fv.solver_parameters();
        fv.explicit_solve();
        return PSST_OK;
    } catch (const std::exception &e) {
        g_case_error = e.what();
        return PSST_HIP_ERROR;
    }
}

void psst_case_close(psst_case *c) { delete c; }

}  // extern "C"
