/* ref_integration_main.cpp -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
 * The reference's main() (main.cpp:9-63) with its explicit_solve() call
 * replaced by the device path exactly as INTEGRATION.md section 2 writes it:
 * the reference reads the mesh and builds its grid, include/psst_ref_adapter.hpp
 * flattens it, libpsst runs the loop, and the reference's save_data() writes
 * the result.  tests/test_ref_adapter.py compiles it against the reference's
 * headers and objects; without a GPU the run stops at psst_ctx_create.
 *
 * usage: psst_ref_integration MESH CASE_PREFIX */
#include <algorithm>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "psstAnsysGridReader.h"
#include "psstFiniteVolumeGrid.h"
#include "psstGridReader.h"
#include "psstOutputManager.h"
#include "psstVariable.h"
#include "psst_ref_adapter.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: psst_ref_integration MESH CASE_PREFIX\n";
        return 2;
    }
    const std::string grid_name = argv[1], output_name = argv[2];
    std::ifstream in_file(grid_name);
    if (!in_file.is_open()) return 1;
    psstInterfaceGridBuilder *builder = new psstAnsysBuilder;
    psstInterfaceGridDirector *director = new psstAnsysDirector;
    director->set_builder(builder);
    psstInterfaceInputGrid *grid = director->construct_grid(in_file);
    in_file.close();
    std::shared_ptr<psstInterfaceOutputManager> output_manager = std::make_shared<psstAnsysOutputManager>(output_name);
    try {
        // ---- INTEGRATION.md section 2, verbatim ----
        psstGridFV* fv = grid->create_finite_volume_grid();
        fv->attach_output_manager(output_manager);
        fv->settings_file_name(output_name);
        fv->solver_parameters();                               // the reference reads its settings and initial field

        PsstRefMesh flat = psst_ref_flatten(*fv);              // face_storage()/volume_storage() -> psst_mesh_desc
        psst_mesh_desc mesh = flat.desc();
        psst_case* cs = nullptr;                               // psst_settings from the same settings files
        if (psst_case_from_mesh(&mesh, (int32_t)flat.zones.size(), flat.zones.data(), output_name.c_str(), &cs))
            throw std::runtime_error(psst_case_last_error());
        psst_settings set; int64_t info[8];
        psst_case_settings(cs, &set);
        psst_case_info(cs, info);                              // dim, N, G, F, stages, first, last, save step
        psst_ctx* ctx = nullptr;
        if (psst_ctx_create(&mesh, &set, &ctx) != PSST_OK) throw std::runtime_error(psst_last_error(nullptr));
        std::vector<double> prim((flat.dim + 3) * (size_t)(flat.N + flat.G));
        psst_ref_read_state(flat, prim.data());                // the reference's own initial field
        auto check = [&](psst_status s) { if (s != PSST_OK) throw std::runtime_error(psst_last_error(ctx)); };
        check(psst_set_state(ctx, prim.data()));
        const int ws = flat.dim + 2;
        for (int64_t it = info[5]; it <= info[6]; ) {          // explicit_solve's loop, a save interval per call
            int32_t n = (int32_t)std::min<int64_t>(info[7] - (it - 1) % info[7], info[6] - it + 1);
            std::vector<double> res((size_t)n * ws);
            check(psst_run(ctx, n, 0.0, res.data(), nullptr));
            // ... print the res rows as psstFiniteVolumeGrid.cpp:369-383 does ...
            it += n;
            check(psst_get_state(ctx, prim.data()));
            psst_ref_write_state(flat, prim.data());           // the device state into the reference's variables
            fv->save_data();                                   // the reference's own writers
        }
        psst_ctx_destroy(ctx);
        psst_case_close(cs);
        // ---- end of the section ----
        delete fv;
    } catch (const std::runtime_error &e) {
        std::cerr << "psst_ref_integration: " << e.what() << "\n";
        return 1;
    }
    delete grid;
    delete director;
    delete builder;
    return 0;
}
