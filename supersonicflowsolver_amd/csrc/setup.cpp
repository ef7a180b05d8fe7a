// setup.cpp -- the reference's settings files and initial state, and its
// output formats.
//
// psstGridFV::solver_parameters (psstFiniteVolumeGrid.cpp:36-194) reads, in
// order: .inp line 1 (init mode) -> .prop + .par or .sav (initial field,
// :641-717) -> .bvp/.btq (boundary conditions and their initial ghost
// states, :403-557) -> the remaining .inp lines.  The same order is kept
// here because the ghost initial states depend on the interior field and the
// reference velocity on the boundary files.
//
// Deliberate differences: a .bvp/.btq ending in a newline is accepted (the
// reference re-reads the last zone with id 0 and dereferences a null
// condition, quirk Q11), and CRLF files are accepted.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "psst_host.hpp"

namespace psst {

namespace {

// psstFiniteVolumeGrid.cpp:20-27: first token of the line, rest of the line
// is a comment.
double get_parameter(std::istream &in) {
    double x = 0.0;
    in >> x;
    while (true) {
        int ch = in.get();
        if (ch == '\n' || ch == EOF) break;
    }
    return x;
}

std::ifstream open_or_throw(const std::string &path, const char *what) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error(std::string("Error: Could not open ") + what + " (" + path + ")");
    return in;
}

int zone_by_name(const Grid &g, const std::string &name) {
    for (size_t i = 0; i < g.zones.size(); ++i)
        if (g.zones[i].name == name) return (int)i;
    throw std::runtime_error("Error: Could not found zone! (" + name + ")");
}

struct PrimView {
    std::vector<double> &p;
    int dim;
    int64_t NT;
    double &u(int j, int64_t c) { return p[(size_t)j * NT + c]; }
    double &P(int64_t c) { return p[(size_t)dim * NT + c]; }
    double &T(int64_t c) { return p[(size_t)(dim + 1) * NT + c]; }
    double &rho(int64_t c) { return p[(size_t)(dim + 2) * NT + c]; }
};

int32_t ghost_neighbour(const Grid &g, int32_t gc) {
    int32_t f = g.ghost_face[(size_t)(gc - g.N)];
    return g.f0[(size_t)f] == gc ? g.f1[(size_t)f] : g.f0[(size_t)f];
}

// Dynamic condition's calculate_at (psstBoundaryCondition.cpp:72-127, :212-219).
void dyn_apply(const Grid &g, const psst_bc &bc, PrimView &q, int32_t gc) {
    int32_t c = ghost_neighbour(g, gc);
    const int dim = g.dim;
    switch (bc.dynamic_kind) {
    case PSST_BC_SS_INLET:
        break;
    case PSST_BC_SS_OUTLET:
        for (int j = 0; j < dim; ++j) q.u(j, gc) = q.u(j, c);
        q.P(gc) = q.P(c);
        break;
    case PSST_BC_SLIP_WALL: {
        int32_t f = g.ghost_face[(size_t)(gc - g.N)];
        double vn = 0.0;
        for (int j = 0; j < dim; ++j) vn += q.u(j, c) * g.fn[(size_t)j * g.F + f];
        for (int j = 0; j < dim; ++j) q.u(j, gc) = q.u(j, c) - vn * g.fn[(size_t)j * g.F + f];
        q.P(gc) = q.P(c);
        break;
    }
    case PSST_BC_OUTLET:
        for (int j = 0; j < dim; ++j) q.u(j, gc) = q.u(j, c);
        break;
    default:
        throw std::runtime_error("bad dynamic boundary condition");
    }
}

void thermal_apply(const Grid &g, const psst_bc &bc, PrimView &q, int32_t gc) {
    int32_t c = ghost_neighbour(g, gc);
    if (bc.thermal_kind == PSST_BCT_OUTLET || bc.thermal_kind == PSST_BCT_WALL_Q) q.T(gc) = q.T(c);
}

}  // namespace

void load_settings(Case &cs, const std::string &prefix) {
    Grid &g = cs.grid;
    const int dim = g.dim;
    const int64_t NT = (int64_t)g.N + g.G;
    cs.prefix = prefix;
    psst_settings &s = cs.settings;
    s = psst_settings{};
    s.precision = 64;
    s.dt_ordering = 0;
    s.fused = 1;
    s.venkat_k = 5.0;
    cs.prim.assign((size_t)(dim + 3) * NT, 0.0);
    PrimView q{cs.prim, dim, NT};

    std::ifstream inp = open_or_throw(prefix + ".inp", "setting file");
    int init_mode = (int)get_parameter(inp);
    // .prop (psstFiniteVolumeGrid.cpp:643-655)
    {
        std::ifstream in = open_or_throw(prefix + ".prop", "fluid properties file");
        cs.M = get_parameter(in);
        s.cp = get_parameter(in);
        double R = 8.31;  // quirk Q10: gas constant hard-coded
        s.rm = R / cs.M;
        s.gamma = s.cp / (s.cp - s.rm);
        s.cv = s.cp / s.gamma;
    }
    if (init_mode == 0) {
        // .par (:658-702)
        cs.iter = 1;
        std::ifstream in = open_or_throw(prefix + ".par", "initial parameters file");
        double vel[3] = {0, 0, 0};
        for (int j = 0; j < dim - 1; ++j) in >> vel[j];
        vel[dim - 1] = get_parameter(in);
        double T = get_parameter(in);
        double p = get_parameter(in);
        for (int32_t c = 0; c < g.N; ++c) {
            for (int j = 0; j < dim; ++j) q.u(j, c) = vel[j];
            q.T(c) = T;
            q.P(c) = p;
            q.rho(c) = p / (s.rm * T);
        }
    } else if (init_mode == 1) {
        // .sav (:559-604)
        std::ifstream in = open_or_throw(prefix + ".sav", "initial parameters file");
        in >> cs.iter;
        while (!in.eof()) {
            int64_t zid = 0;
            if (!(in >> zid) || zid == 0) break;
            int zi = -1;
            for (size_t i = 0; i < g.zones.size(); ++i)
                if (g.zones[i].id == zid) zi = (int)i;
            if (zi < 0 || g.zones[(size_t)zi].bc != 1) throw std::runtime_error("Error! Unknown zone in .sav file");
            const GridZone &z = g.zones[(size_t)zi];
            int64_t n = 0;
            in >> n;
            if (n != z.cell_count) throw std::runtime_error("Error! Different size of storage found!");
            for (int32_t i = 0; i < z.cell_count; ++i) {
                int32_t c = z.cell_begin + i;
                double p, T;
                in >> p >> T;
                q.P(c) = p;
                q.T(c) = T;
                q.rho(c) = p / (s.rm * T);
                for (int j = 0; j < dim; ++j) in >> q.u(j, c);
            }
        }
    } else {
        throw std::runtime_error("Error! Unknown command detected!");
    }
    // .bvp / .btq (:403-557)
    cs.bc.assign(g.zones.size(), psst_bc{});
    cs.ref_vel = 0.0;
    std::vector<bool> has_dyn(g.zones.size(), false), has_th(g.zones.size(), false);
    {
        std::ifstream in = open_or_throw(prefix + ".bvp", "boundary condition file");
        std::string name;
        int64_t id;
        while (in >> name >> id) {
            int zi = zone_by_name(g, name);
            const GridZone &z = g.zones[(size_t)zi];
            if (z.bc == 1 || z.bc == 2 || z.bc == 12 || z.bc == 31 || z.bc == 32)
                throw std::runtime_error("Error: Could not attach boundary condition to inner zone!");
            psst_bc &b = cs.bc[(size_t)zi];
            b.dynamic_kind = (int32_t)id;
            if (id == PSST_BC_SS_INLET) {
                cs.ref_vel = 0.0;
                for (int i = 0; i < dim; ++i) {
                    in >> b.velocity[i];
                    cs.ref_vel += b.velocity[i] * b.velocity[i];
                }
                cs.ref_vel = std::sqrt(cs.ref_vel);
                b.pressure = get_parameter(in);
            } else if (id == PSST_BC_SS_OUTLET || id == PSST_BC_SLIP_WALL) {
            } else if (id == PSST_BC_OUTLET) {
                b.pressure = get_parameter(in);
            } else if (id == PSST_BC_INLET) {
                // the reference's psstInlet constructor always throws (quirk Q9,
                // psstBoundaryCondition.cpp:147); kept as an error here.
                throw std::runtime_error("Error: Wrong dimension! (subsonic inlet 27 is unusable in the reference)");
            } else {
                throw std::runtime_error("Error: unknown dynamic boundary condition id");
            }
            has_dyn[(size_t)zi] = true;
            // initialize() (psstBoundaryCondition.cpp:26-38, :60-64, :96-100, :197-204)
            for (int32_t i = 0; i < z.cell_count; ++i) {
                int32_t gc = z.cell_begin + i;
                if (id == PSST_BC_SS_INLET) {
                    q.P(gc) = b.pressure;
                    for (int j = 0; j < dim; ++j) q.u(j, gc) = b.velocity[j];
                } else if (id == PSST_BC_OUTLET) {
                    q.P(gc) = b.pressure;
                }
                dyn_apply(g, b, q, gc);
            }
        }
    }
    {
        std::ifstream in = open_or_throw(prefix + ".btq", "boundary condition file");
        std::string name;
        int64_t id;
        while (in >> name >> id) {
            int zi = zone_by_name(g, name);
            const GridZone &z = g.zones[(size_t)zi];
            if (z.bc == 1 || z.bc == 2 || z.bc == 12 || z.bc == 31 || z.bc == 32)
                throw std::runtime_error("Error: Could not attach boundary condition to inner zone!");
            psst_bc &b = cs.bc[(size_t)zi];
            b.thermal_kind = (int32_t)id;
            if (id == PSST_BCT_WALL_T || id == PSST_BCT_INLET) {
                b.temperature = get_parameter(in);
                double c = std::sqrt(s.gamma * s.rm * b.temperature);
                if (cs.ref_vel < c) cs.ref_vel = c;
            } else if (id == PSST_BCT_WALL_Q || id == PSST_BCT_OUTLET) {
            } else {
                throw std::runtime_error("Error: unknown thermal boundary condition id");
            }
            has_th[(size_t)zi] = true;
            for (int32_t i = 0; i < z.cell_count; ++i) {
                int32_t gc = z.cell_begin + i;
                if (id == PSST_BCT_WALL_T || id == PSST_BCT_INLET) q.T(gc) = b.temperature;
                thermal_apply(g, b, q, gc);
            }
            for (int32_t i = 0; i < z.cell_count; ++i) {
                int32_t gc = z.cell_begin + i;
                q.rho(gc) = q.P(gc) / (s.rm * q.T(gc));
            }
        }
    }
    for (size_t zi = 0; zi < g.zones.size(); ++zi) {
        const GridZone &z = g.zones[zi];
        if (z.bc == 1 || z.bc == 2 || z.bc == 12 || z.bc == 31 || z.bc == 32) continue;
        if (!has_dyn[zi] || !has_th[zi])
            throw std::runtime_error("Error: boundary zone '" + z.name + "' needs both a .bvp and a .btq entry");
    }
    // remaining .inp lines (:48-192)
    s.cfl = get_parameter(inp);
    s.time_mode = (int32_t)get_parameter(inp);
    if (s.time_mode == PSST_DT_UNIFORM) {
        double rv = cs.ref_vel;
        if (std::abs(rv) < 1.0e-10) rv = 1.0;
        if (dim == 2)
            s.uniform_dt = s.cfl * std::sqrt(g.min_vol) / rv;
        else
            s.uniform_dt = s.cfl * std::pow(g.min_vol, 1.0 / 3.0) / rv;
    } else if (s.time_mode != PSST_DT_LOCAL1 && s.time_mode != PSST_DT_LOCAL2) {
        throw std::runtime_error("Error: unknown time step mode");
    }
    int64_t mi = (int64_t)get_parameter(inp);
    cs.max_iter = mi + cs.iter - 1;
    cs.iter_step = (int64_t)get_parameter(inp);
    if (cs.iter_step <= 0) cs.iter_step = 1;
    s.rk_levels = (int32_t)get_parameter(inp);
    int rk_mode = (int)get_parameter(inp);
    if (rk_mode == 0) {
        // psstRungeKutta::standart_coefficients (psstRungeKutta.cpp:10-49)
        static const double tab[5][5] = {{1.0},
                                         {0.4242, 1.0},
                                         {0.1918, 0.4929, 1.0000},
                                         {0.1084, 0.2602, 0.5052, 1.0000},
                                         {0.0695, 0.1602, 0.2898, 0.5060, 1.0000}};
        if (s.rk_levels < 1 || s.rk_levels > 5) throw std::runtime_error("Error: Unknow RK order!");
        for (int i = 0; i < s.rk_levels; ++i) s.rk_coeff[i] = tab[s.rk_levels - 1][i];
        std::string line;
        std::getline(inp, line);
    } else {
        if (s.rk_levels < 1 || s.rk_levels > 8) throw std::runtime_error("Error: RK level must be 1..8");
        for (int i = 0; i < s.rk_levels - 1; ++i) inp >> s.rk_coeff[i];
        s.rk_coeff[s.rk_levels - 1] = get_parameter(inp);
    }
    s.flux = (int32_t)get_parameter(inp);
    s.order = (int32_t)get_parameter(inp);
    if (s.flux < 1 || s.flux > 3) throw std::runtime_error("Error: Unknow convective flow scheme!");
    if (s.order != 1 && s.order != 2) throw std::runtime_error("Error! Unknown scheme order!");
    int grad = (int)get_parameter(inp);
    s.grad = (grad == 1) ? PSST_GRAD_GREEN_GAUSS : PSST_GRAD_LEAST_SQUARES;
    s.limiter = (int32_t)get_parameter(inp);
    if (s.order == 2 && (s.limiter < 1 || s.limiter > 4))
        throw std::runtime_error("Error: second order needs a limiter (1..4)");
    s.n_bc = (int32_t)cs.bc.size();
    s.bc = cs.bc.data();
}

void write_sav(const Case &c, const double *prim, int64_t iter_value, const std::string &path) {
    // psstFiniteVolumeGrid.cpp:606-639
    const Grid &g = c.grid;
    const int64_t NT = (int64_t)g.N + g.G;
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("Error: Could not found *.sav file!");
    out.precision(11);
    out << iter_value << std::endl;
    for (const GridZone &z : g.zones) {
        if (z.bc != 1) continue;
        out << z.id << std::endl;
        out << z.cell_count << std::endl;
        for (int32_t i = 0; i < z.cell_count; ++i) {
            int64_t cc = z.cell_begin + i;
            out << std::fixed << prim[(size_t)g.dim * NT + cc] << " ";
            out << std::fixed << prim[(size_t)(g.dim + 1) * NT + cc] << " ";
            for (int j = 0; j < g.dim; ++j) out << std::fixed << prim[(size_t)j * NT + cc] << " ";
            out << std::endl;
        }
    }
    out << 0;
}

std::string residual_line(int64_t iter, double seconds, const double *res, int n) {
    // psstFiniteVolumeGrid.cpp:369-382
    std::ostringstream o;
    o << iter << " ";
    o << std::setprecision(11) << std::scientific << seconds << " ";
    for (int j = 0; j < n; ++j) o << std::setprecision(3) << std::scientific << " " << res[j];
    return o.str();
}

}  // namespace psst
