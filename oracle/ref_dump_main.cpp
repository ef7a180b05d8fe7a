/* ref_dump_main.cpp -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
 * A driver of the reference solver compiled from the reference's own sources
 * (read in place under the reference tree).  It does what the reference's
 * main() does (main.cpp:9-63) and then, with `--dump FILE`, writes the final
 * primitive state at full precision (%a), one value per line, in the oracle's
 * canonical order: interior cells (bc 1 zones in ascending id order), then
 * ghost cells (boundary zones in ascending id order); per cell u..., p, T,
 * rho.  Used only in this container to write the golden fixtures under
 * tests/golden/. */
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "psstAnsysGridReader.h"
#include "psstFiniteVolumeGrid.h"
#include "psstGridReader.h"
#include "psstOutputManager.h"
#include "psstVariable.h"

int main(int argc, char **argv) {
    const char *dump = nullptr;
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    std::ifstream in_file("problem_manager");
    if (!in_file.is_open()) { std::cerr << "no problem_manager\n"; return 1; }
    std::string grid_name, output_name;
    in_file >> grid_name >> output_name;
    in_file.close();
    in_file.open(grid_name);
    if (!in_file.is_open()) { std::cerr << "no grid\n"; return 1; }
    psstInterfaceGridBuilder *builder = new psstAnsysBuilder;
    psstInterfaceGridDirector *director = new psstAnsysDirector;
    director->set_builder(builder);
    psstInterfaceInputGrid *grid = director->construct_grid(in_file);
    in_file.close();
    std::shared_ptr<psstInterfaceOutputManager> om = std::make_shared<psstAnsysOutputManager>(output_name);
    psstGridFV *fv = grid->create_finite_volume_grid();
    fv->attach_output_manager(om);
    fv->settings_file_name(output_name);
    fv->solver_parameters();
    fv->explicit_solve();
    if (dump) {
        FILE *f = std::fopen(dump, "w");
        size_t dim = fv->dimension();
        for (int pass = 0; pass < 2; ++pass) {
            for (auto &kv : *fv->volume_storage()) {
                bool interior = kv.second.bc_type_id_ == 1;
                if (pass == 0 && !interior) continue;
                if (pass == 1 && (interior || kv.second.bc_type_id_ == 32)) continue;
                for (auto &vol : kv.second.volumes_) {
                    auto var = vol->variable();
                    for (size_t j = 0; j < dim; ++j) std::fprintf(f, "%a\n", var->velocity(j));
                    std::fprintf(f, "%a\n%a\n%a\n", var->pressure(), var->temperature(), var->density());
                }
            }
        }
        std::fclose(f);
    }
    delete fv;
    delete grid;
    delete director;
    delete builder;
    return 0;
}
