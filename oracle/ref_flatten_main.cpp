/* ref_flatten_main.cpp -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
 * Exercises include/psst_ref_adapter.hpp inside the reference's own build
 * (its sources compiled in place, see the Makefile): the reference reads a
 * Fluent mesh with its director/builder and creates its psstGridFV
 * (main.cpp:43-50), reads its settings files (solver_parameters), and the
 * adapter flattens that grid into a psst_mesh_desc; libpsst then reads the
 * same settings files over the flattened mesh (psst_case_from_mesh).  The
 * driver writes every array of the mesh description, the settings, the
 * library's initial state and the reference's own initial state (its
 * variables after solver_parameters) to a binary file, which
 * tests/test_ref_adapter.py compares with psst_case_open on the same files;
 * then it writes the library's initial state back into the reference's
 * variables and saves it with the reference's own save_data().
 *
 * usage: psst_ref_flatten MESH CASE_PREFIX OUT */
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "psstAnsysGridReader.h"
#include "psstFiniteVolumeGrid.h"
#include "psstGridReader.h"
#include "psstOutputManager.h"
#include "psstVariable.h"
#include "psst_ref_adapter.hpp"

namespace {
// record: 16-byte name, 1-byte kind ('d' double, 'i' int32), int64 count, data
void put(std::FILE *f, const char *name, const double *p, size_t n) {
    char nm[16] = {0};
    std::strncpy(nm, name, 15);
    std::fwrite(nm, 1, 16, f);
    std::fputc('d', f);
    const long long c = (long long)n;
    std::fwrite(&c, sizeof(c), 1, f);
    std::fwrite(p, sizeof(double), n, f);
}
void put(std::FILE *f, const char *name, const int32_t *p, size_t n) {
    char nm[16] = {0};
    std::strncpy(nm, name, 15);
    std::fwrite(nm, 1, 16, f);
    std::fputc('i', f);
    const long long c = (long long)n;
    std::fwrite(&c, sizeof(c), 1, f);
    std::fwrite(p, sizeof(int32_t), n, f);
}
}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::cerr << "usage: psst_ref_flatten MESH CASE_PREFIX OUT\n";
        return 2;
    }
    std::ifstream in_file(argv[1]);
    if (!in_file.is_open()) {
        std::cerr << "cannot open " << argv[1] << "\n";
        return 1;
    }
    psstInterfaceGridBuilder *builder = new psstAnsysBuilder;
    psstInterfaceGridDirector *director = new psstAnsysDirector;
    director->set_builder(builder);
    psstInterfaceInputGrid *grid = director->construct_grid(in_file);
    in_file.close();
    std::shared_ptr<psstInterfaceOutputManager> om = std::make_shared<psstAnsysOutputManager>(argv[2]);
    psstGridFV *fv = grid->create_finite_volume_grid();
    fv->attach_output_manager(om);
    fv->settings_file_name(argv[2]);
    fv->solver_parameters();

    PsstRefMesh m = psst_ref_flatten(*fv);
    psst_mesh_desc d = m.desc();
    psst_case *cs = nullptr;
    if (psst_case_from_mesh(&d, (int32_t)m.zones.size(), m.zones.data(), argv[2], &cs) != PSST_OK) {
        std::cerr << "psst_case_from_mesh: " << psst_case_last_error() << "\n";
        return 1;
    }
    psst_settings s{};
    psst_case_settings(cs, &s);
    const size_t NT = (size_t)m.N + m.G;
    std::vector<double> prim_case((size_t)(m.dim + 3) * NT), prim_ref(prim_case.size());
    psst_case_initial_state(cs, prim_case.data());
    psst_ref_read_state(m, prim_ref.data());

    // the library's initial state back into the reference's variables and
    // out through the reference's own writers (<prefix>.dat/.cas/.sav)
    psst_ref_write_state(m, prim_case.data());
    fv->save_data();

    std::FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    const int32_t sizes[4] = {m.dim, m.N, m.G, m.F};
    put(f, "sizes", sizes, 4);
    put(f, "cell_volume", d.cell_volume, NT);
    put(f, "cell_centre", d.cell_centre, (size_t)m.dim * NT);
    put(f, "face_normal", d.face_normal, (size_t)m.dim * m.F);
    put(f, "face_area", d.face_area, (size_t)m.F);
    put(f, "face_cells", d.face_cells, 2 * (size_t)m.F);
    put(f, "face_beta", d.face_beta, (size_t)m.F);
    put(f, "visit_offsets", d.visit_offsets, (size_t)m.N + 1);
    put(f, "visit_faces", d.visit_faces, (size_t)d.visit_offsets[m.N]);
    put(f, "list_offsets", d.list_offsets, (size_t)m.N + 1);
    put(f, "list_faces", d.list_faces, (size_t)d.list_offsets[m.N]);
    put(f, "ghost_face", d.ghost_face, (size_t)m.G);
    put(f, "ghost_bc", d.ghost_bc, (size_t)m.G);
    put(f, "lsq_r", d.lsq_r, 6 * (size_t)m.N);
    put(f, "lsq_beta", d.lsq_beta, (size_t)m.N);
    put(f, "min_volume", &d.min_volume, 1);
    const int32_t si[6] = {s.flux, s.order, s.grad, s.limiter, s.time_mode, s.rk_levels};
    put(f, "settings_i", si, 6);
    const double sd[14] = {s.rk_coeff[0], s.rk_coeff[1], s.rk_coeff[2], s.rk_coeff[3], s.rk_coeff[4], s.rk_coeff[5],
                           s.rk_coeff[6], s.rk_coeff[7], s.cfl, s.gamma, s.rm, s.cp, s.cv, s.uniform_dt};
    put(f, "settings_d", sd, 14);
    std::vector<double> bc;
    for (int32_t i = 0; i < s.n_bc; ++i) {
        const psst_bc &b = s.bc[i];
        const double row[7] = {(double)b.dynamic_kind, (double)b.thermal_kind, b.velocity[0], b.velocity[1],
                               b.velocity[2], b.pressure, b.temperature};
        bc.insert(bc.end(), row, row + 7);
    }
    put(f, "bc", bc.data(), bc.size());
    put(f, "prim_case", prim_case.data(), prim_case.size());
    put(f, "prim_ref", prim_ref.data(), prim_ref.size());
    std::fclose(f);
    psst_case_close(cs);
    delete fv;
    delete grid;
    delete director;
    delete builder;
    return 0;
}
