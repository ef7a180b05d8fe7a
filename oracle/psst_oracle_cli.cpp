/* psst_oracle_cli.cpp -- TEST INFRASTRUCTURE ONLY.
 * Command-line driver of the CPU oracle that mirrors the reference's main()
 * (main.cpp:9-63) and the output side of psstGridFV::explicit_solve
 * (psstFiniteVolumeGrid.cpp:256-267, :369-400): it reads `problem_manager`
 * in the working directory, runs the case, and writes `<case>_residual.dat`
 * and `<case>.sav` in the reference's formats.  `--dump FILE` also writes
 * the final primitive state as hex floats (%a), one value per line, in the
 * canonical order (interior cells, then ghosts; per cell u..., p, T, rho). */
#include "psst_oracle.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    const char *dump = nullptr;
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    std::ifstream pm("problem_manager");
    if (!pm) { std::cerr << "Error: Could not open problem_manager file!\n"; return 1; }
    std::string grid_name, case_name;
    pm >> grid_name >> case_name;
    void *h = oracle_open(grid_name.c_str(), case_name.c_str());
    if (!h) { std::cerr << "oracle: " << oracle_last_error() << "\n"; return 1; }
    int info[8];
    oracle_info(h, info);
    const int dim = info[0], ws = dim + 2;
    long iter = info[5], max_iter = info[6], iter_step = info[7];
    std::ofstream out;
    if (iter == 1) out.open(case_name + "_residual.dat");
    else out.open(case_name + "_residual.dat", std::ios::app);
    std::vector<double> res(ws);
    auto begin = std::chrono::steady_clock::now();
    for (; iter <= max_iter; ++iter) {
        if (oracle_run(h, 1, res.data())) { std::cerr << "oracle: " << oracle_last_error() << "\n"; return 1; }
        out << iter << " ";
        auto end = std::chrono::steady_clock::now();
        auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(end - begin);
        out << std::setprecision(11) << std::scientific << ms.count() / 1000.0 << " ";
        for (int j = 0; j < ws; ++j) out << std::setprecision(3) << std::scientific << " " << res[j];
        if (iter % iter_step == 0) {
            out << std::endl;
            oracle_write_sav(h, (case_name + ".sav").c_str(), iter);
        } else {
            out << "\n";
        }
    }
    out.close();
    oracle_write_sav(h, (case_name + ".sav").c_str(), iter);
    if (dump) {
        size_t tot = (size_t)info[1] + info[2];
        std::vector<double> prim((size_t)(dim + 3) * tot);
        oracle_get_state(h, prim.data());
        FILE *f = std::fopen(dump, "w");
        for (size_t i = 0; i < tot; ++i)
            for (int j = 0; j < dim + 3; ++j) std::fprintf(f, "%a\n", prim[j * tot + i]);
        std::fclose(f);
    }
    oracle_close(h);
    return 0;
}
