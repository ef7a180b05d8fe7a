// psst_host.hpp -- host-side C++ of the MI355X explicit finite-volume Euler
// solver: Fluent ASCII mesh I/O, the wedge mesh generator, the flattened
// finite-volume grid (geometry and the reference's iteration orders), the
// settings files and the psstGridFV driver that calls the device C-ABI.
//
// Reference surface mirrored here:
//   psstAnsysDirector/psstAnsysBuilder  (psstAnsysGridReader.cpp:948-1374)
//   psstAnsysInputGrid::create_finite_volume_grid (psstAnsysGridReader.cpp:688-907)
//   psstGridFV::{solver_parameters, explicit_solve, save_data}
//                                       (psstFiniteVolumeGrid.h:52-85)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "psst_gpu.h"

namespace psst {
// Phase log of long host set-ups: with PSST_VERBOSE set in the environment,
// prints "psst: <what> (<seconds since the previous line>)" to stderr.
void log_phase(const char *what);
}  // namespace psst

namespace psst {

// ---------------------------------------------------------------------------
// Raw mesh: the content of a Fluent ASCII .msh/.cas file.
// ---------------------------------------------------------------------------
struct RawZone {
    int64_t id = 0;
    int64_t first = 0, last = 0;  // 1-based index range of the zone's faces or cells
    int64_t bc = 0;               // Fluent zone type: 1 fluid cells, 2 interior faces, 31/32 parent
                                  // faces/cells of an adapted mesh, others boundary
    std::string kind;             // type word of the 39/45 section ("wall", "fluid", ...)
    std::string name;             // zone name used by the .bvp/.btq files
    int64_t parent_zone = 0, child_zone = 0;  // links of the 58/59 tree sections
};

struct RawMesh {
    int dim = 0;
    std::vector<double> xyz;         // 3 per vertex (z = 0 in 2D)
    std::vector<int8_t> face_nv;     // vertices per face (2, 3 or 4)
    std::vector<int64_t> face_v;     // 4 per face, 1-based vertex ids
    std::vector<int64_t> face_c;     // 2 per face: cell(0), cell(1) as in the file; 0 = none
    std::vector<int64_t> face_zone;  // zone id per face
    std::vector<int64_t> cell_zone;  // zone id per cell
    std::vector<int8_t> cell_type;   // Fluent element type per cell (1 tri, 2 tet, 3 quad, 4 hex, ...)
    std::vector<RawZone> zones;      // order of first appearance in the file
    // adapted meshes: children of each parent face / cell (1-based ids), in
    // the order the 58/59 sections list them; empty for other meshes
    std::vector<std::vector<int64_t>> face_kids, cell_kids;

    int64_t n_vertices() const { return (int64_t)xyz.size() / 3; }
    int64_t n_faces() const { return (int64_t)face_nv.size(); }
    int64_t n_cells() const { return (int64_t)cell_zone.size(); }
    const RawZone *zone(int64_t id) const;
};

// Fluent ASCII reader: sections 0, 2, 10, 12, 13, 39, 45, 58, 59 in any
// order, LF or CRLF line ends, hex indices.  Adapted meshes keep their parent
// zones (types 31/32) and trees; the solver uses only the active elements.
RawMesh read_fluent_msh(const std::string &path);
// Writer in the layout the reference reader accepts (header lines end with
// "(", LF line ends, vertices as %.20e).
void write_fluent_msh(const RawMesh &mesh, const std::string &path);

// ---------------------------------------------------------------------------
// Wedge generator (compression ramp channel, 2D quads or 3D hex/tet).
// ---------------------------------------------------------------------------
struct WedgeParams {
    int nx = 120, ny = 40, nz = 0;  // nz = 0: 2D
    double length = 3.0, height = 1.0, depth = 1.0;
    double ramp_start = 0.5, ramp_length = 0.6, ramp_deg = 15.0;
    double jitter = 0.0;            // fraction of local spacing (+-)
    uint64_t jitter_seed = 1;
    int numbering = 0;              // 0 column-major (ix*ny+iy), 1 RCM, 2 random permutation
    uint64_t perm_seed = 2;
    int tets = 0;                   // 3D: split each hex into 6 tetrahedra
    int periodic_z = 0;             // 3D: z faces as one "side" zone pair
};
RawMesh generate_wedge(const WedgeParams &p);

// ---------------------------------------------------------------------------
// Flattened finite-volume grid in canonical numbering:
//   cells   0..N-1   interior cells, zone (bc 1) ascending id, then zone order
//           N..N+G-1 ghost cells, boundary zones ascending id, then zone order
//   faces   the reference's face-loop order: face zones ascending id (bc 31
//           skipped), then zone order
// ---------------------------------------------------------------------------
struct GridZone {
    int64_t id = 0;
    int64_t bc = 0;
    std::string name;
    int32_t cell_begin = 0, cell_count = 0;  // canonical cell range (volume zones)
    int32_t face_begin = 0, face_count = 0;  // canonical face range (face zones)
    bool has_cells = false, has_faces = false;
};

struct Grid {
    int dim = 2;
    int32_t N = 0, G = 0, F = 0;
    std::vector<double> vol;         // [N+G]
    std::vector<double> cx;          // [3][N+G]
    std::vector<double> fn;          // [3][F] unit normal (points toward cell 0)
    std::vector<double> fa;          // [F]
    std::vector<double> fc;          // [3][F] face centre
    std::vector<int32_t> f0, f1;     // [F] neighbour 0 / 1
    std::vector<int32_t> fcell;      // [2][F] = f0 then f1
    std::vector<int32_t> visit_off, visit_face;  // [N+1], faces per interior cell in face-loop order
    std::vector<int32_t> list_off, list_face;    // [N+G+1], faces per cell in face-list order
    std::vector<double> beta;        // [F] Green-Gauss weight (reference indexing, quirk Q1)
    std::vector<double> lsq_r;       // [6][N] least-squares factors (quirk Q20 indexing)
    std::vector<double> lsq_beta;    // [N]
    std::vector<int32_t> ghost_face; // [G]
    std::vector<int32_t> ghost_zone; // [G] index into zones
    std::vector<int32_t> cell_zone_local;  // [N] zone-local index of interior cell
    std::vector<GridZone> zones;     // ascending id
    double min_vol = -1.0;
    std::vector<int64_t> original_cell;   // [N] 1-based file cell index
    std::vector<double> fshift;      // [3][F] translation of cell 1 into cell 0's frame (periodic faces)
    bool periodic = false;           // the mesh has periodic (type 12) face zones
};

Grid build_grid(const RawMesh &mesh);

// ---------------------------------------------------------------------------
// Settings (.inp/.prop/.par/.bvp/.btq/.sav) and the initial state.
// ---------------------------------------------------------------------------
struct Case {
    Grid grid;
    psst_settings settings{};
    std::vector<psst_bc> bc;          // one per grid zone (index = GridZone index)
    std::vector<double> prim;         // [(dim+3)][N+G] initial primitive state
    double M = 0, ref_vel = 0;
    int64_t iter = 1, max_iter = 0, iter_step = 1;
    std::string prefix;
};

void load_settings(Case &c, const std::string &prefix);

// ---------------------------------------------------------------------------
// Output formats of the reference.
// ---------------------------------------------------------------------------
// `.sav` checkpoint (psstFiniteVolumeGrid.cpp:606-639)
void write_sav(const Case &c, const double *prim, int64_t iter_value, const std::string &path);
// one line of `<case>_residual.dat` (psstFiniteVolumeGrid.cpp:369-396)
std::string residual_line(int64_t iter, double seconds, const double *res, int n);
// `.dat` field file of a primitive state [(dim+3)][N+G] (psstOutputManager.cpp:71-128)
void write_dat(const Grid &g, const double *prim, const std::string &path);
// `.cas` mesh file (psstOutputManager.cpp:130-398)
void write_cas(const RawMesh &m, const std::string &path);

// ---------------------------------------------------------------------------
// Mesh descriptor view (pointers into a Grid) for psst_ctx_create.
// ---------------------------------------------------------------------------
psst_mesh_desc make_mesh_desc(const Grid &g);

}  // namespace psst
