// output.cpp -- the reference's field and mesh output files, written from a
// primitive state and the mesh as read:
//   * `<case>.dat`: every primitive variable of every interior and ghost
//     cell in Fluent's "(300 ...)" data-section layout
//     (psstAnsysOutputManager::private_write_field and
//     private_write_variable_to_tecplot, psstOutputManager.cpp:71-128);
//   * `<case>.cas`: the mesh in Fluent's case layout: dimension, vertices,
//     faces and cells per zone, the cell and face trees of an adapted mesh,
//     zone names (private_write_cas_file and the writers it calls,
//     psstOutputManager.cpp:130-398).
// The reference writes both from its volume and face storages, which are
// std::maps keyed by zone id; here that is ascending zone id over the mesh's
// zones.  Number formats (decimal/hexadecimal switches, precision, spacing,
// blank lines) are the reference's, so the files compare byte for byte.
#include <algorithm>
#include <fstream>
#include <iomanip>
#include <map>
#include <stdexcept>

#include "psst_host.hpp"

namespace psst {

namespace {

// bc_types_names_ (psstAnsysOutputManager::initialize, psstOutputManager.cpp:8-27)
const char *zone_type_name(int64_t bc) {
    switch (bc) {
    case 1: return "fluid";
    case 2: return "interior";
    case 3: return "wall";
    case 4: return "pressure-inlet";
    case 5: return "pressure-outlet";
    case 7: return "symmetry";
    case 8: return "periodic-shadow";
    case 9: return "pressure-far-field";
    case 10: return "velocity-inlet";
    case 12: return "periodic";
    case 14: return "radiator";
    case 20: return "mass-flow-inlet";
    case 24: return "interface";
    case 31:
    case 32: return "parent";
    case 36: return "outflow";
    case 37: return "axis";
    default: throw std::runtime_error("Error: no zone type name for zone type " + std::to_string(bc));
    }
}

// Zones that carry data, ascending id (the reference's std::map order).
std::vector<const RawZone *> zones_by_id(const RawMesh &m) {
    std::vector<const RawZone *> z;
    for (const RawZone &r : m.zones)
        if (!(r.first == 0 && r.last == 0)) z.push_back(&r);
    std::sort(z.begin(), z.end(), [](const RawZone *a, const RawZone *b) { return a->id < b->id; });
    return z;
}

bool is_cell_zone(int64_t bc) { return bc == 1 || bc == 32; }

std::ofstream open_append(const std::string &path) {
    std::ofstream out(path, std::ios::app);
    if (!out.is_open()) throw std::runtime_error("Error: Could not open output file " + path);
    return out;
}

}  // namespace

void write_dat(const Grid &g, const double *prim, const std::string &path) {
    {
        std::ofstream trunc(path);
        if (!trunc.is_open()) throw std::runtime_error("Error: Could not open output *.dat file!");
    }
    const int64_t NT = (int64_t)g.N + g.G;
    // variable ids of the Fluent data sections (ansys_variable_data_id,
    // psstOutputManager.h:28-36) and the primitive row each one reads
    std::vector<std::pair<int, int>> vars = {{1, g.dim}, {3, g.dim + 1}, {101, g.dim + 2}, {111, 0}, {112, 1}};
    if (g.dim == 3) vars.push_back({113, 2});
    for (const auto &var : vars) {
        std::ofstream out = open_append(path);
        out.precision(11);
        size_t first = 0, last = 0;
        // volume zones: interior cell zones and the ghost zones of boundary
        // face zones, parent cell zones left out (:84-104)
        for (const GridZone &z : g.zones) {
            if (!z.has_cells) continue;
            first = last;
            last += (size_t)z.cell_count;
            out << "(\t " << 300 << " \t(\t " << var.first << " \t " << z.id << " \t " << 1 << " \t " << 0 << " \t "
                << 0 << " \t " << first + 1 << " \t " << last << " \t) " << std::endl
                << "(\n";
            const double *row = prim + (size_t)var.second * NT;
            for (int32_t j = 0; j < z.cell_count; ++j) out << std::fixed << row[z.cell_begin + j] << "\n";
            out << ")\n" << std::endl;
        }
    }
}

void write_cas(const RawMesh &m, const std::string &path) {
    {
        std::ofstream trunc(path);
        if (!trunc.is_open()) throw std::runtime_error("Error: Could not open output *.cas file!");
    }
    const std::vector<const RawZone *> zones = zones_by_id(m);
    const int dim = m.dim;
    {   // private_write_dimension (:130-143)
        std::ofstream out = open_append(path);
        out << "(0 \"Dimension:\")\n";
        out << "(" << 2 << " " << dim << ")\n" << std::endl;
    }
    {   // private_write_vertices (:145-175)
        std::ofstream out = open_append(path);
        const int64_t nv = m.n_vertices();
        out << "(0 \"Vertices:\")\n";
        out << "(" << std::dec << 10 << " (" << 0 << " " << std::hex << 1 << " " << nv << " " << 1 << " " << dim
            << "))\n";
        out << "(" << std::dec << 10 << " (" << 1 << " " << std::hex << 1 << " " << nv << " " << 1 << " " << dim
            << ")(\n";
        for (int64_t i = 0; i < nv; ++i) {
            for (int j = 0; j < dim; ++j)
                out << std::dec << std::setprecision(20) << std::fixed << std::left << std::scientific << " "
                    << m.xyz[3 * (size_t)i + j] << " ";
            out << "\n";
        }
        out << "))\n" << std::endl;
    }
    {   // private_write_faces (:177-224): every face zone, parents included;
        // a face's neighbours are the file's cell ids, 0 for a ghost or none
        std::ofstream out = open_append(path);
        out << "(0 \"Faces:\")\n";
        out << "(" << std::dec << 13 << " (" << std::hex << 0 << " " << 1 << " " << m.n_faces() << " " << 0 << "))\n";
        for (const RawZone *z : zones) {
            if (is_cell_zone(z->bc)) continue;
            out << "(" << std::dec << 13 << " (" << std::hex << z->id << " " << z->first << " " << z->last << " "
                << z->bc << " " << 0 << ")(\n";
            for (int64_t i = z->first; i <= z->last; ++i) {
                const size_t fi = (size_t)i - 1;
                const int nvf = m.face_nv[fi];
                out << std::hex << std::left << " " << nvf << " ";
                for (int j = 0; j < nvf; ++j) out << std::hex << std::left << " " << m.face_v[4 * fi + j] << " ";
                for (int k = 0; k < 2; ++k) out << std::hex << std::left << " " << m.face_c[2 * fi + k] << " ";
                out << std::endl;
            }
            out << "))\n" << std::endl;
        }
    }
    {   // private_write_cells (:226-260): fluid and parent cell zones, the
        // element type of each cell, nine to a line
        std::ofstream out = open_append(path);
        out << "(0 \"Cells:\")\n";
        out << "(" << std::dec << 12 << " (" << std::hex << 0 << " " << 1 << " " << m.n_cells() << " " << 0 << "))\n";
        for (const RawZone *z : zones) {
            if (!is_cell_zone(z->bc)) continue;
            out << "(" << std::dec << 12 << " (" << std::hex << z->id << " " << z->first << " " << z->last << " "
                << z->bc << " " << 0 << ")(\n";
            for (int64_t i = 0; i <= z->last - z->first; ++i) {
                if (i % 9 == 0 && i != 0) out << std::endl;
                const size_t ci = (size_t)(z->first + i - 1);
                out << " " << (m.cell_type.empty() ? 0 : (int)m.cell_type[ci]);
            }
            out << std::endl << "))\n" << std::endl;
        }
    }
    // private_write_cell_tree / private_write_face_tree (:262-346): one tree
    // section per parent zone, each parent's children as file ids
    auto write_tree = [&](int64_t parent_bc, int section, const char *title,
                          const std::vector<std::vector<int64_t>> &kids) {
        std::ofstream out = open_append(path);
        bool any = false;
        for (const RawZone *z : zones) any = any || z->bc == parent_bc;
        if (!any) return;
        out << "(0 \"" << title << "\")\n";
        for (const RawZone *z : zones) {
            if (z->bc != parent_bc) continue;
            out << "(" << std::dec << section << " (" << std::hex << z->first << " " << z->last << " " << z->id << " "
                << z->child_zone << ")(\n";
            for (int64_t i = z->first; i <= z->last; ++i) {
                static const std::vector<int64_t> none;
                const std::vector<int64_t> &k = (size_t)i <= kids.size() ? kids[(size_t)i - 1] : none;
                out << " " << k.size();
                for (int64_t c : k) out << " " << c;
                out << "\n";
            }
            out << "))\n" << std::endl;
        }
    };
    write_tree(32, 58, "Cells Trees:", m.cell_kids);
    write_tree(31, 59, "Faces Trees:", m.face_kids);
    {   // private_write_zones (:348-377): fluid cell zones, then every face
        // zone but the parents
        std::ofstream out = open_append(path);
        out << "(0 \"Zones:\")\n";
        for (const RawZone *z : zones) {
            if (z->bc != 1) continue;
            out << "(" << std::dec << 39 << " (" << z->id << " " << zone_type_name(z->bc) << " " << z->name << ")())\n"
                << std::endl;
        }
        for (const RawZone *z : zones) {
            if (is_cell_zone(z->bc) || z->bc == 31) continue;
            out << "(" << std::dec << 39 << " (" << z->id << " " << zone_type_name(z->bc) << " " << z->name << ")())\n"
                << std::endl;
        }
    }
}

}  // namespace psst
