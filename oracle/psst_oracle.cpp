/*
 * psst_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial CPU restatement of the reference solver's explicit finite-volume
 * Euler path (PSST, psstGridFV::explicit_solve and the strategy objects it
 * drives).  It is the checker for the HIP path: only tests/, the package's
 * smoke() and bench.py's CPU-baseline leg load it, and never as the thing
 * measured or shipped.  The product path in supersonicflowsolver_amd/ does
 * not link, load or call anything in this directory.
 *
 * It follows the reference operation by operation (same loop orders, same
 * association of every floating-point expression, no FMA contraction) so
 * that its FP64 results are bit-identical to a build of the reference.
 * Every function names the reference lines it restates.  The data structures
 * are flat arrays instead of the reference's shared_ptr object graph.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - whole-run fields: bit-identical to a g++ build of the reference
 *     (oracle/Makefile target `ref`) on the committed fixtures in
 *     tests/golden/ (80x240 HLL/RK1/lc=1 reference-held run; generated
 *     40x120 wedge runs; a small matrix of schemes; the reference's 3D hex
 *     meshes; its adapted meshes with parent zones and face/cell trees);
 *   - `.sav` and residual file of the reference-held run byte-compared;
 *   - limiters, Roe average, fluxes, geometry, gradients also pinned by
 *     closed forms and invariants.
 * 3D least squares (reference psstGradCalculation.cpp:336-339 writes
 * alpha[3] out of bounds) has no reference behaviour to follow; the oracle
 * implements the textbook 3D QR form and it is pinned by linear exactness
 * only.
 */
#include "psst_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

/* ------------------------------------------------------------------------- */
/* Raw mesh (Fluent ASCII sections 2/10/12/13/39/45).                         */
/* ------------------------------------------------------------------------- */

struct RawZone {
    long id = 0;
    long first = 0, last = 0;
    long bc = 0;
    std::string name;
};

struct RawFace {
    int nv = 0;
    long v[4] = {0, 0, 0, 0};
    long c[2] = {0, 0}; /* c[0] = first cell index in the file, c[1] = second */
    long zone = 0;
};

struct RawMesh {
    int dim = 0;
    std::vector<double> xyz; /* 3 per vertex */
    std::vector<RawFace> faces;
    std::vector<long> cell_zone;
    std::vector<RawZone> zones;       /* file order of first appearance */
    std::map<long, size_t> zone_pos;  /* id -> position in zones */
};

struct Cursor {
    const std::string &s;
    size_t p = 0;
    explicit Cursor(const std::string &str) : s(str) {}
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\t' || s[p] == '\r')) ++p;
    }
    bool eat(char ch) {
        ws();
        if (p < s.size() && s[p] == ch) { ++p; return true; }
        return false;
    }
    void expect(char ch) {
        if (!eat(ch)) throw std::runtime_error(std::string("mesh parse: expected '") + ch + "'");
    }
    std::string token() {
        ws();
        size_t b = p;
        while (p < s.size() && !(s[p] == ' ' || s[p] == '\n' || s[p] == '\t' || s[p] == '\r' ||
                                 s[p] == '(' || s[p] == ')'))
            ++p;
        return s.substr(b, p - b);
    }
    long hex() {
        std::string t = token();
        if (t.empty()) throw std::runtime_error("mesh parse: expected hex integer");
        return std::stol(t, nullptr, 16);
    }
    long dec() {
        std::string t = token();
        if (t.empty()) throw std::runtime_error("mesh parse: expected integer");
        return std::stol(t, nullptr, 10);
    }
    double real() {
        std::string t = token();
        if (t.empty()) throw std::runtime_error("mesh parse: expected real");
        return std::strtod(t.c_str(), nullptr);
    }
    std::vector<long> header_hex() { /* "( a b c ... )" -> hex values */
        expect('(');
        std::vector<long> v;
        while (true) {
            ws();
            if (p < s.size() && s[p] == ')') { ++p; break; }
            v.push_back(hex());
        }
        return v;
    }
    void skip_to_close() { /* skip to the ')' closing the current level */
        int depth = 1;
        while (p < s.size() && depth > 0) {
            if (s[p] == '"') {
                ++p;
                while (p < s.size() && s[p] != '"') ++p;
            } else if (s[p] == '(') {
                ++depth;
            } else if (s[p] == ')') {
                --depth;
            }
            ++p;
        }
    }
};

void add_zone(RawMesh &m, long id, long first, long last, long bc) {
    /* psstAnsysGridReader.cpp:600-616 add_zones_storage(id, first, last, bc) */
    auto it = m.zone_pos.find(id);
    if (it == m.zone_pos.end()) {
        RawZone z;
        z.id = id; z.first = first; z.last = last; z.bc = bc;
        m.zone_pos[id] = m.zones.size();
        m.zones.push_back(z);
    } else {
        RawZone &z = m.zones[it->second];
        z.first = first; z.last = last; z.bc = bc;
    }
}

void add_zone_name(RawMesh &m, long id, const std::string &name) {
    /* psstAnsysGridReader.cpp:618-632 add_zones_storage(id, name) */
    auto it = m.zone_pos.find(id);
    if (it == m.zone_pos.end()) {
        RawZone z;
        z.id = id; z.name = name;
        m.zone_pos[id] = m.zones.size();
        m.zones.push_back(z);
    } else {
        m.zones[it->second].name = name;
    }
}

void add_zone_connection(RawMesh &m, long parent, long first, long last, long child) {
    /* psstAnsysGridReader.cpp:655-686 add_zones_storage_connection: a parent
     * zone seen first here is created as a parent face zone (type 31) with
     * the tree's range; a child zone seen first here gets an empty storage.
     * Both keep that position in the zone order when their own section
     * arrives later and overwrites range and type. */
    if (m.zone_pos.find(parent) == m.zone_pos.end()) {
        RawZone z;
        z.id = parent; z.first = first; z.last = last; z.bc = 31;
        m.zone_pos[parent] = m.zones.size();
        m.zones.push_back(z);
    }
    if (m.zone_pos.find(child) == m.zone_pos.end()) {
        RawZone z;
        z.id = child;
        m.zone_pos[child] = m.zones.size();
        m.zones.push_back(z);
    }
}

RawMesh read_mesh(const std::string &path) {
    /* Semantics of psstAnsysDirector::build_step and psstAnsysBuilder::read_*
     * (psstAnsysGridReader.cpp:948-1116, :1138-1306); a plain recursive
     * parser, tolerant of CRLF and of section order. */
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open mesh " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    RawMesh m;
    Cursor c(text);
    int nd_hint = 0;
    while (true) {
        c.ws();
        if (c.p >= text.size()) break;
        if (text[c.p] != '(') { ++c.p; continue; }
        ++c.p;
        long sec = c.dec();
        if (sec == 2) {
            m.dim = (int)c.dec();
            c.skip_to_close();
        } else if (sec == 10) {
            std::vector<long> h = c.header_hex();
            if (h.size() < 3) throw std::runtime_error("bad node section header");
            long zone = h[0], first = h[1], last = h[2];
            long nd = h.size() > 4 ? h[4] : 0;
            if (zone == 0) { nd_hint = (int)nd; c.skip_to_close(); continue; }
            int dim = m.dim ? m.dim : (int)nd;
            c.expect('(');
            long n = last - first + 1;
            if ((long)m.xyz.size() / 3 != first - 1) throw std::runtime_error("vertex sections out of order");
            for (long i = 0; i < n; ++i) {
                double x[3] = {0, 0, 0};
                for (int k = 0; k < dim; ++k) x[k] = c.real();
                m.xyz.insert(m.xyz.end(), x, x + 3);
            }
            c.expect(')');
            c.skip_to_close();
        } else if (sec == 12) {
            std::vector<long> h = c.header_hex();
            if (h.size() < 3) throw std::runtime_error("bad cell section header");
            long zone = h[0], first = h[1], last = h[2];
            if (zone == 0) { c.skip_to_close(); continue; }
            if (h.size() < 5) throw std::runtime_error("bad cell section header");
            long bc = h[3], type = h[4];
            add_zone(m, zone, first, last, bc);
            if ((long)m.cell_zone.size() < last) m.cell_zone.resize(last, 0);
            for (long i = first; i <= last; ++i) m.cell_zone[i - 1] = zone;
            if (c.eat('(')) {
                if (type == 0)
                    for (long i = first; i <= last; ++i) c.hex();
                c.expect(')');
            }
            c.skip_to_close();
        } else if (sec == 13) {
            std::vector<long> h = c.header_hex();
            if (h.size() < 3) throw std::runtime_error("bad face section header");
            long zone = h[0], first = h[1], last = h[2];
            if (zone == 0) { c.skip_to_close(); continue; }
            if (h.size() < 5) throw std::runtime_error("bad face section header");
            long bc = h[3], ftype = h[4];
            add_zone(m, zone, first, last, bc);
            c.expect('(');
            if ((long)m.faces.size() != first - 1) throw std::runtime_error("face sections out of order");
            for (long i = first; i <= last; ++i) {
                RawFace rf;
                long t = ftype;
                if (ftype == 0) t = c.hex();
                if (t < 2 || t > 4) throw std::runtime_error("unsupported face type");
                rf.nv = (int)t;
                for (int k = 0; k < rf.nv; ++k) rf.v[k] = c.hex();
                rf.c[0] = c.hex();
                rf.c[1] = c.hex();
                rf.zone = zone;
                m.faces.push_back(rf);
            }
            c.expect(')');
            c.skip_to_close();
        } else if (sec == 39 || sec == 45) {
            c.expect('(');
            long id = c.dec();
            c.token();
            std::string name = c.token();
            c.skip_to_close();
            c.skip_to_close();
            add_zone_name(m, id, name);
        } else if (sec == 58 || sec == 59) {
            /* read_cell_tree / read_face_tree (:1068-1116): the parent-child
             * links themselves are not used by the solver; only the zone
             * storages they create (add_zones_storage_connection, :655-686)
             * matter, through the file order of zones */
            std::vector<long> h = c.header_hex();
            if (h.size() < 4) throw std::runtime_error("bad tree section header");
            add_zone_connection(m, h[2], h[0], h[1], h[3]);
            c.skip_to_close();
        } else {
            c.skip_to_close();
        }
    }
    if (m.dim == 0) m.dim = nd_hint;
    if (m.dim != 2 && m.dim != 3) throw std::runtime_error("mesh dimension missing");
    return m;
}

/* ------------------------------------------------------------------------- */
/* Finite-volume grid (psstAnsysInputGrid::create_finite_volume_grid).        */
/* ------------------------------------------------------------------------- */

struct Face {
    long zone;
    bool boundary;
    int nv;
    long vtx[4];
    int nb[2];
    double n[3] = {0, 0, 0};
    double A = 0.0;
    double c[3] = {0, 0, 0};
    /* periodic faces (zone type 12, no reference counterpart): translation
     * carrying neighbour 1 into neighbour 0's frame */
    double sh[3] = {0, 0, 0};
    bool periodic = false;
    double own_beta = 0.0; /* periodic faces: their own Green-Gauss weight */
};

struct Vol {
    long zone;
    bool boundary;
    std::vector<int> faces; /* face-list order (psstVolume::add_face order) */
    std::vector<int> nbs;   /* neighbour across faces[i] */
    double V = 0.0;
    double x[3] = {0, 0, 0};
    /* state (psstVariable) */
    double u[3] = {0, 0, 0}, T = 0, p = 0, rho = 0;
    double g[6][3]; /* grads: u_0..u_{d-1}, p, T, rho  (index d, d+1, d+2) */
    double w[5] = {0, 0, 0, 0, 0}, w0[5] = {0, 0, 0, 0, 0}, R[5] = {0, 0, 0, 0, 0};
    double dt = 0.0;
    Vol() { std::memset(g, 0, sizeof(g)); }
};

struct ZoneStore {
    long bc = 0;
    std::string name;
    std::vector<int> items;
};

struct BCDyn {
    int kind = 0; /* 23, 26, 15, 27, 25 */
    double u[3] = {0, 0, 0};
    double p = 0;
};
struct BCThermal {
    int kind = 0; /* 11, 18, 27, 25 */
    double T = 0;
};

struct Oracle {
    int dim = 2;
    std::vector<double> xyz;
    std::vector<Face> faces;
    std::vector<Vol> vols;
    std::map<long, ZoneStore> face_zones; /* std::map: ascending zone id */
    std::map<long, ZoneStore> vol_zones;
    std::vector<long> file_zone_order;
    double min_vol = -1.0;
    double venkat_k = 5.0, venkat_eps2 = 0.0;  // limiter 4 (no reference counterpart)
    size_t n_faces_total = 0, n_cells_total = 0;

    /* settings */
    double M = 0, cp = 0, rm = 0, gamma = 0, cv = 0, ref_vel = 0;
    double cfl = 0.5;
    int lc = 1, rk_level = 1, flux = 2, order = 1, grad = 1, limiter = 3;
    std::vector<double> rk;
    long iter = 1, max_iter = 0, iter_step = 1;
    std::map<long, BCDyn> dyn;
    std::map<long, BCThermal> thermal;
    double uniform_dt = 0.0;

    /* Green-Gauss beta table indexed by zone-local face index (quirk Q1). */
    std::vector<double> gg_beta;
    /* Least-squares factors indexed by zone-local cell index (quirk Q20). */
    std::vector<std::array<double, 6>> lsq_r;
    std::vector<double> lsq_beta;

    bool w_ready = false;
    std::string err;
};

void face_initialize(Oracle &o, Face &f) {
    /* psstFace::initialize, psstFiniteVolumeGridTools.cpp:397-505 */
    const int dim = o.dim;
    auto V = [&](int j, int k) { return o.xyz[3 * (f.vtx[j] - 1) + k]; };
    f.A = 0.0;
    for (int i = 0; i < 3; ++i) { f.n[i] = 0.0; f.c[i] = 0.0; }
    if (dim == 2) {
        if (f.nv != 2) throw std::runtime_error("2D face with !=2 vertices");
        for (int i = 0; i < 2; ++i) {
            f.c[i] = 0.5 * (V(1, i) + V(0, i));
            double comp = V(1, i) - V(0, i);
            f.A += comp * comp;
        }
        f.A = std::sqrt(f.A);
        f.n[0] = V(0, 1) - V(1, 1);
        f.n[1] = V(1, 0) - V(0, 0);
    } else {
        int nv = f.nv;
        double avg[3] = {0, 0, 0}, avg3[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < nv; ++j) avg[i] += V(j, i);
            avg[i] /= (double)(size_t)nv;
            avg3[i] = avg[i] * 3;
        }
        double a[3], b[3], cc[3];
        double ts;
        int last = nv - 1;
        for (int i = 0; i < last; ++i) {
            for (int j = 0; j < 3; ++j) {
                a[j] = V(i, j) - avg[j];
                b[j] = V(i + 1, j) - avg[j];
            }
            cc[0] = a[1] * b[2] - a[2] * b[1];
            cc[1] = a[2] * b[0] - a[0] * b[2];
            cc[2] = a[0] * b[1] - a[1] * b[0];
            ts = 0.0;
            for (int j = 0; j < 3; ++j) ts += cc[j] * cc[j];
            ts = std::sqrt(ts);
            f.A += ts;
            for (int j = 0; j < 3; ++j) f.c[j] += ts * (a[j] + b[j] + avg3[j]);
        }
        for (int j = 0; j < 3; ++j) {
            a[j] = V(last, j) - avg[j];
            b[j] = V(0, j) - avg[j];
        }
        cc[0] = a[1] * b[2] - a[2] * b[1];
        cc[1] = a[2] * b[0] - a[0] * b[2];
        cc[2] = a[0] * b[1] - a[1] * b[0];
        ts = 0.0;
        for (int j = 0; j < 3; ++j) ts += cc[j] * cc[j];
        ts = std::sqrt(ts);
        f.A += ts;
        double sq3 = f.A * 3;
        for (int j = 0; j < 3; ++j) {
            f.c[j] += ts * (a[j] + b[j] + avg3[j]);
            f.c[j] /= sq3;
            f.n[j] = cc[j];
        }
        f.A *= 0.5;
    }
    /* psstGeometryVector::initialize + normalize (:248-256, :279-290) */
    double len = 0.0;
    for (int i = 0; i < dim; ++i) len += f.n[i] * f.n[i];
    len = std::sqrt(len);
    if (std::abs(len) < 1.0e-10) throw std::runtime_error("degenerate face normal");
    for (int i = 0; i < dim; ++i) f.n[i] /= len;
}

void volume_initialize(Oracle &o, int vi) {
    /* psstVolume::initialize, psstFiniteVolumeGridTools.cpp:709-827 */
    Vol &v = o.vols[vi];
    const int dim = o.dim;
    v.V = 0.0;
    for (int i = 0; i < 3; ++i) v.x[i] = 0.0;
    size_t nf = v.faces.size();
    v.nbs.resize(nf);
    for (size_t i = 0; i < nf; ++i) {
        const Face &f = o.faces[v.faces[i]];
        if (f.nb[0] != vi && f.nb[1] != vi) throw std::runtime_error("face/volume link broken");
        v.nbs[i] = (f.nb[0] == vi) ? f.nb[1] : f.nb[0];
    }
    if (v.boundary) {
        const Face &f = o.faces[v.faces.at(0)];
        for (int i = 0; i < 3; ++i) v.x[i] = f.c[i];
        return;
    }
    /* a periodic face as this volume sees it: translated by -sh for neighbour 1 */
    auto OFF = [&](const Face &f, int k) { return (f.periodic && f.nb[1] == vi) ? -f.sh[k] : 0.0; };
    auto VX = [&](const Face &f, int j, int k) {
        const double x = o.xyz[3 * (f.vtx[j] - 1) + k], t = OFF(f, k);
        return t != 0.0 ? x + t : x;
    };
    auto FCc = [&](const Face &f, int k) {
        const double t = OFF(f, k);
        return t != 0.0 ? f.c[k] + t : f.c[k];
    };
    if (dim == 2) {
        double avg[2] = {0, 0}, avg3[2] = {0, 0};
        for (int j = 0; j < 2; ++j) {
            for (size_t i = 0; i < nf; ++i) avg[j] += o.faces[v.faces[i]].c[j];
            avg[j] /= (double)nf;
            avg3[j] = avg[j] * 3;
        }
        double a[2], b[2];
        for (size_t i = 0; i < nf; ++i) {
            const Face &f = o.faces[v.faces[i]];
            for (int k = 0; k < 2; ++k) {
                a[k] = VX(f, 0, k) - avg[k];
                b[k] = VX(f, 1, k) - avg[k];
            }
            double tv = std::abs(a[0] * b[1] - a[1] * b[0]);
            v.V += tv;
            for (int k = 0; k < 2; ++k) v.x[k] += tv * (a[k] + b[k] + avg3[k]);
        }
        double v3 = v.V * 3;
        for (int k = 0; k < 2; ++k) v.x[k] /= v3;
        v.V /= 2;
    } else {
        double avg[3] = {0, 0, 0}, avg4[3] = {0, 0, 0};
        for (int j = 0; j < 3; ++j) {
            for (size_t i = 0; i < nf; ++i) avg[j] += FCc(o.faces[v.faces[i]], j);
            avg[j] /= (double)nf;
            avg4[j] = avg[j] * 4;
        }
        double a[3], b[3], c[3];
        for (size_t i = 0; i < nf; ++i) {
            const Face &f = o.faces[v.faces[i]];
            for (int j = 0; j < f.nv; ++j) {
                int next = j + 1;
                if (next == f.nv) next = 0;
                for (int k = 0; k < 3; ++k) {
                    a[k] = VX(f, j, k) - avg[k];
                    b[k] = VX(f, next, k) - avg[k];
                    c[k] = FCc(f, k) - avg[k];
                }
                double tv = std::abs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                                     a[2] * (b[0] * c[1] - b[1] * c[0]));
                v.V += tv;
                for (int k = 0; k < 3; ++k) v.x[k] += tv * (a[k] + b[k] + c[k] + avg4[k]);
            }
        }
        double v4 = v.V * 4.0;
        for (int k = 0; k < 3; ++k) v.x[k] /= v4;
        v.V /= 6.0;
    }
}

void build_grid(Oracle &o, const RawMesh &m) {
    /* psstAnsysGridReader.cpp:688-907 */
    o.dim = m.dim;
    o.xyz = m.xyz;
    o.n_faces_total = m.faces.size();
    o.n_cells_total = m.cell_zone.size();
    /* storage allocation in file zone order (:712-783) */
    std::map<long, std::vector<int>> vol_of_zone, face_of_zone;
    for (const RawZone &z : m.zones) {
        o.file_zone_order.push_back(z.id);
        long n = z.last - z.first + 1;
        if (z.bc == 1 || z.bc == 32) {
            ZoneStore &zs = o.vol_zones[z.id];
            zs.bc = z.bc; zs.name = z.name;
            for (long j = 0; j < n; ++j) {
                Vol v; v.zone = z.id; v.boundary = false;
                zs.items.push_back((int)o.vols.size());
                o.vols.push_back(v);
            }
        } else if (z.bc == 2 || z.bc == 31 || z.bc == 12) {
            ZoneStore &fs = o.face_zones[z.id];
            fs.bc = z.bc; fs.name = z.name;
            for (long j = 0; j < n; ++j) {
                const RawFace &rf = m.faces[z.first + j - 1];
                Face f; f.zone = z.id; f.boundary = false; f.nv = rf.nv;
                for (int k = 0; k < 4; ++k) f.vtx[k] = rf.v[k];
                f.nb[0] = f.nb[1] = -1;
                fs.items.push_back((int)o.faces.size());
                o.faces.push_back(f);
            }
        } else {
            ZoneStore &vs = o.vol_zones[z.id];
            vs.bc = z.bc; vs.name = z.name;
            ZoneStore &fs = o.face_zones[z.id];
            fs.bc = z.bc; fs.name = z.name;
            for (long j = 0; j < n; ++j) {
                const RawFace &rf = m.faces[z.first + j - 1];
                Face f; f.zone = z.id; f.boundary = true; f.nv = rf.nv;
                for (int k = 0; k < 4; ++k) f.vtx[k] = rf.v[k];
                f.nb[0] = f.nb[1] = -1;
                fs.items.push_back((int)o.faces.size());
                o.faces.push_back(f);
                Vol v; v.zone = z.id; v.boundary = true;
                vs.items.push_back((int)o.vols.size());
                o.vols.push_back(v);
            }
        }
    }
    /* face-volume links in file zone order (:825-883) */
    for (const RawZone &z : m.zones) {
        if (z.bc == 1 || z.bc == 32) continue;
        ZoneStore &fs = o.face_zones.at(z.id);
        for (size_t j = 0; j < fs.items.size(); ++j) {
            const RawFace &rf = m.faces[z.first + j - 1];
            Face &f = o.faces[fs.items[j]];
            for (int k = 0; k < 2; ++k) {
                long cell = rf.c[k];
                int vi;
                if (z.bc == 31) {
                    /* parent faces (:860-874): neighbours linked, but the
                     * face joins no cell's face list and no ghost exists */
                    if (cell == 0) {
                        f.nb[k] = -1;
                        continue;
                    }
                    long vz = m.cell_zone.at(cell - 1);
                    const RawZone &cz = m.zones[m.zone_pos.at(vz)];
                    f.nb[k] = o.vol_zones.at(vz).items.at(cell - cz.first);
                    continue;
                }
                if (cell == 0) {
                    vi = o.vol_zones.at(z.id).items.at(j);
                } else {
                    long vz = m.cell_zone.at(cell - 1);
                    const RawZone &cz = m.zones[m.zone_pos.at(vz)];
                    vi = o.vol_zones.at(vz).items.at(cell - cz.first);
                }
                f.nb[k] = vi;
                o.vols[vi].faces.push_back(fs.items[j]);
            }
            face_initialize(o, f);
        }
    }
    /* periodic zones: periodic along z with period P = the z extent of the
     * vertices; a face's vertices lie in neighbour 0's frame, neighbour 1 sees
     * the face translated by +P (sh = -P) */
    {
        bool any = false;
        for (const RawZone &z : m.zones) any = any || z.bc == 12;
        if (any) {
            if (o.dim != 3) throw std::runtime_error("periodic zones need a 3D mesh");
            double zmin = 1e300, zmax = -1e300;
            for (size_t v = 0; v < o.xyz.size() / 3; ++v) {
                zmin = std::min(zmin, o.xyz[3 * v + 2]);
                zmax = std::max(zmax, o.xyz[3 * v + 2]);
            }
            for (const RawZone &z : m.zones) {
                if (z.bc != 12) continue;
                for (int fi : o.face_zones.at(z.id).items) {
                    o.faces[fi].periodic = true;
                    o.faces[fi].sh[2] = zmin - zmax;
                }
            }
        }
    }
    /* volume initialisation in file zone order, V_min over bc==1 (:885-905) */
    double mn = -1.0;
    for (const RawZone &z : m.zones) {
        if (z.bc == 2 || z.bc == 31 || z.bc == 32 || z.bc == 12) continue;
        for (int vi : o.vol_zones.at(z.id).items) {
            volume_initialize(o, vi);
            if (z.bc == 1) {
                if (mn > o.vols[vi].V || mn < 0.0) mn = o.vols[vi].V;
            }
        }
    }
    o.min_vol = mn;
}

/* ------------------------------------------------------------------------- */
/* Settings files (psstFiniteVolumeGrid.cpp:20-194, :403-717).                */
/* ------------------------------------------------------------------------- */

double get_parameter(std::ifstream &in) {
    /* psstFiniteVolumeGrid.cpp:20-27 */
    double x = 0.0;
    in >> x;
    while (in.get() != '\n' && !in.eof()) continue;
    return x;
}

void initialize_field(Oracle &o, const std::string &prefix, int mode) {
    /* private_initialize_field, psstFiniteVolumeGrid.cpp:641-717 */
    std::ifstream in(prefix + ".prop");
    if (!in) throw std::runtime_error("cannot open " + prefix + ".prop");
    o.M = get_parameter(in);
    o.cp = get_parameter(in);
    double R = 8.31;
    o.rm = R / o.M;
    o.gamma = o.cp / (o.cp - o.rm);
    o.cv = o.cp / o.gamma;
    in.close();
    if (mode == 0) {
        o.iter = 1;
        std::ifstream par(prefix + ".par");
        if (!par) throw std::runtime_error("cannot open " + prefix + ".par");
        double vel[3] = {0, 0, 0};
        for (int j = 0; j < o.dim - 1; ++j) par >> vel[j];
        vel[o.dim - 1] = get_parameter(par);
        double T = get_parameter(par);
        double p = get_parameter(par);
        for (auto &kv : o.vol_zones) {
            if (kv.second.bc != 1) continue;
            for (int vi : kv.second.items) {
                Vol &v = o.vols[vi];
                for (int j = 0; j < o.dim; ++j) v.u[j] = vel[j];
                v.T = T;
                v.p = p;
                v.rho = p / (o.rm * T);
            }
        }
    } else if (mode == 1) {
        /* private_load_parameter, :559-604 */
        std::ifstream sav(prefix + ".sav");
        if (!sav) throw std::runtime_error("cannot open " + prefix + ".sav");
        sav >> o.iter;
        while (!sav.eof()) {
            long zid = 0;
            sav >> zid;
            if (zid == 0) break;
            ZoneStore &zs = o.vol_zones.at(zid);
            size_t n = 0;
            sav >> n;
            if (n != zs.items.size()) throw std::runtime_error("Different size of storage found");
            for (size_t i = 0; i < n; ++i) {
                Vol &v = o.vols[zs.items[i]];
                double p, T;
                sav >> p; v.p = p;
                sav >> T; v.T = T;
                v.rho = p / (o.rm * T);
                for (int j = 0; j < o.dim; ++j) sav >> v.u[j];
            }
        }
    } else {
        throw std::runtime_error("unknown init mode");
    }
}

ZoneStore &find_zone_by_name(Oracle &o, const std::string &name, long &id) {
    for (auto &kv : o.vol_zones) {
        if (kv.second.name == name) { id = kv.first; return kv.second; }
    }
    throw std::runtime_error("Could not found zone " + name);
}

void dyn_calculate_at(Oracle &o, const BCDyn &bc, Vol &g);
void thermal_calculate_at(Oracle &o, const BCThermal &bc, Vol &g);

void load_boundary(Oracle &o, const std::string &prefix) {
    /* private_load_boundary, psstFiniteVolumeGrid.cpp:403-557.  A line that
     * fails to parse ends the file (the reference's while(!eof) would re-read
     * the last zone with bc_id 0 and crash, quirk Q11). */
    std::ifstream in(prefix + ".bvp");
    if (!in) throw std::runtime_error("cannot open " + prefix + ".bvp");
    while (!in.eof()) {
        std::string name;
        long bc_id = 0;
        if (!(in >> name >> bc_id)) break;
        long zid;
        ZoneStore &zs = find_zone_by_name(o, name, zid);
        if (zs.bc == 1 || zs.bc == 2 || zs.bc == 12 || zs.bc == 31 || zs.bc == 32)
            throw std::runtime_error("Could not attach boundary condition to inner zone");
        BCDyn d;
        d.kind = (int)bc_id;
        if (bc_id == 23) {
            o.ref_vel = 0.0;
            for (int i = 0; i < o.dim; ++i) {
                in >> d.u[i];
                o.ref_vel += d.u[i] * d.u[i];
            }
            o.ref_vel = std::sqrt(o.ref_vel);
            d.p = get_parameter(in);
        } else if (bc_id == 26 || bc_id == 15) {
        } else if (bc_id == 27) {
            throw std::runtime_error("subsonic inlet (27) always throws in the reference (Q9)");
        } else if (bc_id == 25) {
            d.p = get_parameter(in);
        } else {
            throw std::runtime_error("unknown dynamic boundary condition");
        }
        o.dyn[zid] = d;
        /* initialize() of the dynamic condition (psstBoundaryCondition.cpp) */
        for (int gi : zs.items) {
            Vol &g = o.vols[gi];
            if (d.kind == 23) {
                g.p = d.p;
                for (int j = 0; j < o.dim; ++j) g.u[j] = d.u[j];
            } else if (d.kind == 25) {
                g.p = d.p;
            }
            dyn_calculate_at(o, d, g);
        }
    }
    std::ifstream tq(prefix + ".btq");
    if (!tq) throw std::runtime_error("cannot open " + prefix + ".btq");
    while (!tq.eof()) {
        std::string name;
        long bc_id = 0;
        if (!(tq >> name >> bc_id)) break;
        long zid;
        ZoneStore &zs = find_zone_by_name(o, name, zid);
        if (zs.bc == 1 || zs.bc == 2 || zs.bc == 12 || zs.bc == 31 || zs.bc == 32)
            throw std::runtime_error("Could not attach boundary condition to inner zone");
        BCThermal t;
        t.kind = (int)bc_id;
        if (bc_id == 11 || bc_id == 27) {
            t.T = get_parameter(tq);
            double c = std::sqrt(o.gamma * o.rm * t.T);
            if (o.ref_vel < c) o.ref_vel = c;
        } else if (bc_id == 18 || bc_id == 25) {
        } else {
            throw std::runtime_error("unknown thermal boundary condition");
        }
        o.thermal[zid] = t;
        for (int gi : zs.items) {
            Vol &g = o.vols[gi];
            if (t.kind == 11 || t.kind == 27) g.T = t.T;
            thermal_calculate_at(o, t, g);
        }
        for (int gi : zs.items) {
            Vol &g = o.vols[gi];
            g.rho = g.p / (o.rm * g.T);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Boundary conditions (psstBoundaryCondition.cpp).                           */
/* ------------------------------------------------------------------------- */

void dyn_calculate_at(Oracle &o, const BCDyn &bc, Vol &g) {
    const Vol &nb = o.vols[g.nbs.at(0)];
    const int dim = o.dim;
    switch (bc.kind) {
    case 23: /* psstSupersonicInlet: calculate_at is a no-op (.h:52) */
        break;
    case 26: /* psstSupersonicOutlet::calculate_at :72-80 */
        for (int j = 0; j < dim; ++j) g.u[j] = nb.u[j];
        g.p = nb.p;
        break;
    case 15: { /* psstInviscidWall::calculate_at :108-127 */
        const Face &f = o.faces[g.faces.at(0)];
        double vn = 0.0;
        for (int j = 0; j < dim; ++j) vn += nb.u[j] * f.n[j];
        for (int j = 0; j < dim; ++j) g.u[j] = nb.u[j] - vn * f.n[j];
        g.p = nb.p;
        break;
    }
    case 25: /* psstOutlet::calculate_at :212-219 */
        for (int j = 0; j < dim; ++j) g.u[j] = nb.u[j];
        break;
    default:
        throw std::runtime_error("bad dynamic BC");
    }
}

void thermal_calculate_at(Oracle &o, const BCThermal &bc, Vol &g) {
    const Vol &nb = o.vols[g.nbs.at(0)];
    switch (bc.kind) {
    case 27: /* psstThermalInlet: no-op (.h:181) */
    case 11: /* psstThermalWall_T: no-op (.h:231) */
        break;
    case 25: /* psstThermalOutlet :285-288 */
    case 18: /* psstThermalWall_Q :354-357 */
        g.T = nb.T;
        break;
    default:
        throw std::runtime_error("bad thermal BC");
    }
}

/* ------------------------------------------------------------------------- */
/* Limiters (psstScalarLimiters.cpp) and Roe average (psstConvectiveFlow.cpp).*/
/* ------------------------------------------------------------------------- */

double limiter_value(int kind, double num, double den, double eps2 = 0.0) {
    const double eps = 1.0e-10;
    if ((num > 0.0 && den < 0.0) || (num < 0.0 && den > 0.0)) return 0.0;
    switch (kind) {
    case 4: { /* Venkatakrishnan, ratio form (no reference counterpart):
                 (a^2 + 2ab + eps2) / (a^2 + ab + 2b^2 + eps2) */
        if (std::abs(den) < eps) return 1.0;
        const double nn = num * num, nd = num * den, dd = den * den;
        return (nn + 2.0 * nd + eps2) / (nn + nd + 2.0 * dd + eps2);
    }
    case 1: { /* VanAlbada :17-31 */
        if (std::abs(den) < eps) return 1.0;
        double r = num / den;
        return (r * r + r) / (r * r + 1.0);
    }
    case 2: { /* MinMod :46-62 */
        if (std::abs(den) < eps) return 1.0;
        double r = num / den;
        double res = r;
        if (r > 1.0) res = 1.0;
        return res;
    }
    case 3: { /* VanLeer :78-93 */
        if (std::abs(den) < eps) return 2.0;
        double r = num / den;
        return 2.0 * r / (1.0 + r);
    }
    default:
        throw std::runtime_error("limiter not attached");
    }
}

inline double average_by_roe(double lv, double rv, double ld, double rd) {
    /* psstConvectiveFlow.cpp:13-24 */
    double l = std::sqrt(ld);
    double r = std::sqrt(rd);
    double num = lv * l + rv * r;
    double den = (l + r);
    return num / den;
}

/* ------------------------------------------------------------------------- */
/* Gradients (psstGradCalculation.cpp).                                       */
/* ------------------------------------------------------------------------- */

void gg_initialize(Oracle &o) {
    /* psstGreenGauss::initialize :17-72 -- beta indexed by zone-local face
     * index i; later zones overwrite earlier ones (quirk Q1). */
    o.gg_beta.assign(o.n_faces_total, 0.0);
    for (auto &kv : o.face_zones) {
        if (kv.second.bc == 31) continue;
        const auto &items = kv.second.items;
        for (size_t i = 0; i < items.size(); ++i) {
            const Face &f = o.faces[items[i]];
            double d[2];
            for (int j = 0; j < 2; ++j) {
                d[j] = 0.0;
                const Vol &v = o.vols[f.nb[j]];
                if (!v.boundary) {
                    double dist = 0.0;
                    for (int k = 0; k < o.dim; ++k) {
                        double xc = v.x[k];
                        if (j == 1 && f.periodic) xc = xc + f.sh[k];
                        double add = (f.c[k] - xc);
                        add = add * add;
                        dist += add;
                    }
                    d[j] = std::sqrt(dist);
                }
            }
            o.gg_beta[i] = d[1] / (d[0] + d[1]);
            if (f.periodic) o.faces[items[i]].own_beta = d[1] / (d[0] + d[1]);
        }
    }
}

inline double *var_ptr(Oracle &o, Vol &v, int n) {
    if (n < o.dim) return &v.u[n];
    if (n == o.dim) return &v.p;
    if (n == o.dim + 1) return &v.T;
    return &v.rho;
}

void gg_calculate(Oracle &o) {
    /* psstGreenGauss::calculate :80-177 */
    const int dim = o.dim;
    for (auto &kv : o.face_zones) {
        if (kv.second.bc == 31) continue;
        for (int fi : kv.second.items) {
            const Face &f = o.faces[fi];
            for (int j = 0; j < 2; ++j) std::memset(o.vols[f.nb[j]].g, 0, sizeof(o.vols[0].g));
        }
    }
    for (auto &kv : o.face_zones) {
        if (kv.second.bc == 31) continue;
        const auto &items = kv.second.items;
        for (size_t i = 0; i < items.size(); ++i) {
            const Face &f = o.faces[items[i]];
            Vol *v[2] = {&o.vols[f.nb[0]], &o.vols[f.nb[1]]};
            double geom[2];
            for (int j = 0; j < 2; ++j) geom[j] = f.A / v[j]->V;
            double beta = f.periodic ? f.own_beta : o.gg_beta[i];
            double rbeta = 1 - beta;
            for (int n = 0; n < dim + 3; ++n) {
                double tv = beta * *var_ptr(o, *v[0], n) + rbeta * *var_ptr(o, *v[1], n);
                int sign = 1;
                for (int j = 0; j < 2; ++j) {
                    sign *= -1;
                    for (int k = 0; k < dim; ++k) v[j]->g[n][k] += sign * tv * geom[j] * f.n[k];
                }
            }
        }
    }
}

void lsq_initialize(Oracle &o) {
    /* psstLeastSquare::initialize :191-260 (a,b,d,c,e,g = 0..5), indexed by
     * zone-local cell index (quirk Q20). */
    const int dim = o.dim;
    o.lsq_r.assign(o.n_cells_total, std::array<double, 6>{0, 0, 0, 0, 0, 0});
    o.lsq_beta.assign(o.n_cells_total, 0.0);
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc != 1) continue;
        const auto &items = kv.second.items;
        for (size_t i = 0; i < items.size(); ++i) {
            const Vol &v = o.vols[items[i]];
            double cf[6] = {0, 0, 0, 0, 0, 0};
            for (size_t j = 0; j < v.faces.size(); ++j) {
                const Vol &nb = o.vols[v.nbs[j]];
                const Face &fj = o.faces[v.faces[j]];
                double diff[3], w = 0.0;
                for (int k = 0; k < dim; ++k) {
                    double xn = nb.x[k];
                    if (fj.periodic && fj.sh[k] != 0.0) xn = (fj.nb[0] == items[i]) ? xn + fj.sh[k] : xn - fj.sh[k];
                    diff[k] = xn - v.x[k];
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
            auto &r = o.lsq_r[i];
            r[0] = std::sqrt(cf[0]);
            r[1] = cf[1] / r[0];
            r[2] = std::sqrt(cf[2] - r[1] * r[1]);
            o.lsq_beta[i] = 0.0;
            if (dim == 3) {
                r[3] = cf[3] / r[0];
                r[4] = (cf[4] - r[1] / r[0] * cf[3]) / r[2];
                r[5] = std::sqrt(cf[5] - (r[3] * r[3] + r[4] * r[4]));
                o.lsq_beta[i] = (r[1] * r[4] - r[3] * r[2]) / (r[0] * r[2]);
            }
        }
    }
}

void lsq_calculate(Oracle &o) {
    /* psstLeastSquare::calculate :268-368.  2D follows the reference exactly;
     * 3D uses the textbook QR weights (the reference's 3D branch writes out
     * of bounds and is undefined). */
    const int dim = o.dim;
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc != 1) continue;
        const auto &items = kv.second.items;
        for (size_t i = 0; i < items.size(); ++i) {
            Vol &v = o.vols[items[i]];
            std::memset(v.g, 0, sizeof(v.g));
            const auto &r = o.lsq_r[i];
            for (size_t j = 0; j < v.faces.size(); ++j) {
                Vol &nb = o.vols[v.nbs[j]];
                const Face &fj = o.faces[v.faces[j]];
                double diff[3], w = 0.0, alpha[3] = {0, 0, 0}, theta[3] = {0, 0, 0}, phi[6];
                for (int k = 0; k < dim; ++k) {
                    double xn = nb.x[k];
                    if (fj.periodic && fj.sh[k] != 0.0) xn = (fj.nb[0] == items[i]) ? xn + fj.sh[k] : xn - fj.sh[k];
                    diff[k] = xn - v.x[k];
                    w += diff[k] * diff[k];
                }
                for (int n = 0; n < dim; ++n) phi[n] = (nb.u[n] - v.u[n]);
                phi[dim] = (nb.p - v.p);
                phi[dim + 1] = (nb.T - v.T);
                phi[dim + 2] = (nb.rho - v.rho);
                double ws = 1.0 / w;
                if (dim == 2) {
                    alpha[0] = diff[0] / (r[0] * r[0]);
                    alpha[1] = (diff[1] - r[1] / r[0] * diff[0]) / (r[2] * r[2]);
                    theta[0] = alpha[0] - r[1] / r[0] * alpha[1];
                    theta[1] = alpha[1];
                } else {
                    double beta = o.lsq_beta[i];
                    alpha[0] = diff[0] / (r[0] * r[0]);
                    alpha[1] = (diff[1] - r[1] / r[0] * diff[0]) / (r[2] * r[2]);
                    alpha[2] = (diff[2] - r[4] / r[2] * diff[1] + beta * diff[0]) / (r[5] * r[5]);
                    theta[0] = alpha[0] - r[1] / r[0] * alpha[1] + beta * alpha[2];
                    theta[1] = alpha[1] - r[4] / r[2] * alpha[2];
                    theta[2] = alpha[2];
                }
                for (int n = 0; n < dim + 3; ++n)
                    for (int k = 0; k < dim; ++k) v.g[n][k] += ws * theta[k] * phi[n];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Reconstruction (psstValueReconstructor.cpp).                               */
/* ------------------------------------------------------------------------- */

struct RVar {
    double u[3], p, T, rho;
};

void reconstruct(Oracle &o, const Face &f, RVar rv[2]) {
    const int dim = o.dim;
    const Vol *v[2] = {&o.vols[f.nb[0]], &o.vols[f.nb[1]]};
    for (int j = 0; j < 2; ++j) {
        rv[j].p = v[j]->p;
        rv[j].T = v[j]->T;
        rv[j].rho = v[j]->rho;
        for (int i = 0; i < dim; ++i) rv[j].u[i] = v[j]->u[i];
    }
    if (o.order == 1) return; /* psstFirstOrderReconstructor :28-48 */
    /* psstReconstructorDM::reconstruct_vals :70-134 */
    double RLR[3];
    for (int j = 0; j < dim; ++j) RLR[j] = v[0]->x[j] - (f.periodic ? v[1]->x[j] + f.sh[j] : v[1]->x[j]);
    for (int i = 0; i < dim + 2; ++i) {
        double *p0 = (i < dim) ? &rv[0].u[i] : (i == dim ? &rv[0].p : &rv[0].T);
        double *p1 = (i < dim) ? &rv[1].u[i] : (i == dim ? &rv[1].p : &rv[1].T);
        const double *g0 = v[0]->g[i];
        const double *g1 = v[1]->g[i];
        double dl = 0.0, dr = 0.0;
        for (int k = 0; k < dim; ++k) {
            dl += g1[k] * RLR[k];
            dr += g0[k] * RLR[k];
        }
        double vlm = *p1 - (2.0 * dl - (*p0 - *p1));
        double vlp = *p0;
        double vrm = *p1;
        double vrp = *p0 + (2.0 * dr - (*p0 - *p1));
        double nl = vlp - *p1;
        double dnl = *p1 - vlm;
        double nr = vrm - *p0;
        double dnr = *p0 - vrp;
        *p0 += 0.5 * limiter_value(o.limiter, nr, dnr, o.venkat_eps2) * dnr;
        *p1 += 0.5 * limiter_value(o.limiter, nl, dnl, o.venkat_eps2) * dnl;
    }
    for (int j = 0; j < 2; ++j) rv[j].rho = rv[j].p / (o.rm * rv[j].T);
}

/* ------------------------------------------------------------------------- */
/* Convective fluxes (psstConvectiveFlow.cpp).                                */
/* ------------------------------------------------------------------------- */

void boundary_flux(Oracle &o, const Face &f) {
    /* boundary-zone branch shared by HLL/HLLC/AUSM, e.g. :123-204 (Q5) */
    const int dim = o.dim;
    Vol *v[2] = {&o.vols[f.nb[0]], &o.vols[f.nb[1]]};
    int gj = v[0]->boundary ? 0 : 1;
    const Vol &g = *v[gj];
    double vn = 0.0, vsq = 0.0;
    double p = g.p, T = g.T, rho = g.rho, vel[3];
    for (int k = 0; k < dim; ++k) {
        vel[k] = g.u[k];
        vn += vel[k] * f.n[k];
        vsq += vel[k] * vel[k];
    }
    double H = o.cp * T + 0.5 * vsq;
    double flow[5];
    for (int k = 0; k < dim; ++k) flow[k] = rho * vel[k] * vn + p * f.n[k];
    flow[dim] = rho * vn;
    flow[dim + 1] = rho * vn * H;
    if (gj == 0) {
        for (int k = 0; k < dim + 2; ++k) v[1]->R[k] += flow[k] * f.A;
    } else {
        for (int k = 0; k < dim + 2; ++k) v[0]->R[k] -= flow[k] * f.A;
    }
}

struct Gas {
    double gamma, rm, cp, cv;
};

/* Interior-face Riemann flux of two (reconstructed) states; rv[0] is the
 * state of neighbour 0 (the side the normal points to), rv[1] of neighbour 1.
 * HLL psstConvectiveFlow.cpp:215-286, HLLC :499-585, AUSM :782-851. */
void riemann_flux(int flux, int dim, const Gas &gas, const double *n, const RVar rv[2], double *flow) {
    const int ws = dim + 2;
    double p[2], T[2], rho[2], vel[2][3], vn[2], a[2], H[2], E[2], M[2];
    double fi[2][5], wi[2][5];
    for (int j = 0; j < 2; ++j) {
        vn[j] = 0.0;
        double vsq = 0.0;
        p[j] = rv[j].p;
        T[j] = rv[j].T;
        rho[j] = rv[j].rho;
        for (int k = 0; k < dim; ++k) {
            vel[j][k] = rv[j].u[k];
            vn[j] += vel[j][k] * n[k];
            vsq += vel[j][k] * vel[j][k];
        }
        a[j] = std::sqrt(gas.gamma * gas.rm * T[j]);
        H[j] = gas.cp * T[j] + 0.5 * vsq;
        if (flux == 1) {
            M[j] = vn[j] / a[j];
        } else {
            E[j] = gas.cv * T[j] + 0.5 * vsq;
            for (int k = 0; k < dim; ++k) {
                fi[j][k] = rho[j] * vel[j][k] * vn[j] + p[j] * n[k];
                wi[j][k] = rho[j] * vel[j][k];
            }
            wi[j][dim] = rho[j];
            wi[j][dim + 1] = rho[j] * E[j];
            fi[j][dim] = rho[j] * vn[j];
            fi[j][dim + 1] = rho[j] * vn[j] * H[j];
        }
    }
    if (flux == 2 || flux == 3) {
        double Hr = average_by_roe(H[1], H[0], rho[1], rho[0]);
        double vrn = 0.0, vsq = 0.0, vr[3];
        for (int k = 0; k < dim; ++k) {
            vr[k] = average_by_roe(vel[1][k], vel[0][k], rho[1], rho[0]);
            vrn += vr[k] * n[k];
            vsq += vr[k] * vr[k];
        }
        double ar = std::sqrt((gas.gamma - 1.0) * (Hr - 0.5 * vsq));
        double s[2], temp;
        s[1] = vn[1] - a[1];
        temp = vrn - ar;
        if (s[1] > temp) s[1] = temp;
        s[0] = vn[0] + a[0];
        temp = vrn + ar;
        if (s[0] < temp) s[0] = temp;
        if (flux == 2) {
            if (s[1] >= 0.0) {
                for (int j = 0; j < ws; ++j) flow[j] = fi[1][j];
            } else if (s[1] <= 0.0 && s[0] >= 0.0) {
                for (int j = 0; j < ws; ++j)
                    flow[j] = (s[0] * fi[1][j] - s[1] * fi[0][j] + s[0] * s[1] * (wi[0][j] - wi[1][j])) / (s[0] - s[1]);
            } else if (s[0] <= 0.0) {
                for (int j = 0; j < ws; ++j) flow[j] = fi[0][j];
            } else {
                throw std::runtime_error("Could not calculate convective HLL flow");
            }
        } else {
            double sh = (p[0] - p[1] + rho[1] * vn[1] * (s[1] - vn[1]) - rho[0] * vn[0] * (s[0] - vn[0])) /
                        (rho[1] * (s[1] - vn[1]) - rho[0] * (s[0] - vn[0]));
            if (s[1] >= 0.0) {
                for (int j = 0; j < ws; ++j) flow[j] = fi[1][j];
            } else if ((s[1] <= 0.0 && sh >= 0.0) || (sh <= 0.0 && s[0] >= 0.0)) {
                int K = (s[1] <= 0.0 && sh >= 0.0) ? 1 : 0;
                temp = s[K] - sh;
                for (int j = 0; j < dim; ++j)
                    flow[j] = (sh * (s[K] * wi[K][j] - fi[K][j]) +
                               s[K] * (p[K] + rho[K] * (s[K] - vn[K]) * (sh - vn[K])) * n[j]) / temp;
                flow[dim] = (sh * (s[K] * wi[K][dim] - fi[K][dim])) / temp;
                flow[dim + 1] = (sh * (s[K] * wi[K][dim + 1] - fi[K][dim + 1]) +
                                 s[K] * (p[K] + rho[K] * (s[K] - vn[K]) * (sh - vn[K])) * sh) / temp;
            } else if (s[0] <= 0.0) {
                for (int j = 0; j < ws; ++j) flow[j] = fi[0][j];
            } else {
                throw std::runtime_error("Could not calculate convective HLLC flow");
            }
        }
    } else {
        double Ms[2], ps[2];
        if (M[1] >= 1.0) { Ms[1] = M[1]; ps[1] = p[1]; }
        else if (M[1] <= -1.0) { Ms[1] = 0.0; ps[1] = 0.0; }
        else {
            Ms[1] = 0.25 * (M[1] + 1.0) * (M[1] + 1.0);
            ps[1] = 0.25 * p[1] * (M[1] + 1.0) * (M[1] + 1.0) * (2.0 - M[1]);
        }
        if (M[0] >= 1.0) { Ms[0] = 0.0; ps[0] = 0.0; }
        else if (M[0] <= -1.0) { Ms[0] = M[0]; ps[0] = p[0]; }
        else {
            Ms[0] = -0.25 * (M[0] - 1.0) * (M[0] - 1.0);
            ps[0] = 0.25 * p[0] * (M[0] - 1.0) * (M[0] - 1.0) * (2.0 + M[0]);
        }
        double Mf = Ms[0] + Ms[1];
        double pf = ps[0] + ps[1];
        int K = (Mf > 0.0) ? 1 : 0;
        double temp = Mf * a[K] * rho[K];
        for (int j = 0; j < dim; ++j) {
            double pflow = pf * n[j];
            flow[j] = temp * vel[K][j] + pflow;
        }
        flow[dim] = temp;
        flow[dim + 1] = temp * H[K];
    }
}

void interior_flux(Oracle &o, const Face &f) {
    const int ws = o.dim + 2;
    RVar rv[2];
    reconstruct(o, f, rv);
    Vol *v[2] = {&o.vols[f.nb[0]], &o.vols[f.nb[1]]};
    double flow[5];
    Gas gas{o.gamma, o.rm, o.cp, o.cv};
    riemann_flux(o.flux, o.dim, gas, f.n, rv, flow);
    int sign = 1;
    for (int j = 0; j < 2; ++j) {
        sign = (-1) * sign;
        for (int k = 0; k < ws; ++k) v[j]->R[k] += sign * flow[k] * f.A;
    }
}

void convective_calculate(Oracle &o) {
    /* psstConvective*::calculate face loop: map order, skip bc 31 */
    for (auto &kv : o.face_zones) {
        if (kv.second.bc == 31) continue;
        if (kv.second.bc != 2 && kv.second.bc != 12) {
            for (int fi : kv.second.items) boundary_flux(o, o.faces[fi]);
        } else {
            for (int fi : kv.second.items) interior_flux(o, o.faces[fi]);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Local time step (psstLocalTimeCalculation.cpp).                            */
/* ------------------------------------------------------------------------- */

void time_step_at(Oracle &o, Vol &v) {
    const int dim = o.dim;
    if (o.lc == 0) return; /* psstUniformTime::calculate_at is a no-op (.h:41) */
    double c = std::sqrt(o.gamma * o.rm * v.T);
    double den = 0.0;
    if (o.lc == 1) {
        /* psstLocalAccTime1::calculate_at :90-131 (Q3: normal_neighbour is
         * never reset between faces) */
        double nn = 0.0;
        for (size_t j = 0; j < v.faces.size(); ++j) {
            double nv = 0.0;
            const Face &f = o.faces[v.faces[j]];
            const Vol &nb = o.vols[v.nbs[j]];
            double cn = (std::sqrt(o.gamma * o.rm * nb.T) + c) * 0.5;
            for (int k = 0; k < dim; ++k) {
                double t = f.n[k];
                nv += v.u[k] * t;
                nn += nb.u[k] * t;
            }
            nn = (nn + nv) * 0.5;
            den += (std::abs(nn) + cn) * f.A;
        }
    } else {
        /* psstLocalAccTime2::calculate_at :175-213 */
        double s[3] = {0, 0, 0};
        for (size_t j = 0; j < v.faces.size(); ++j) {
            const Face &f = o.faces[v.faces[j]];
            for (int k = 0; k < dim; ++k) s[k] += f.A * std::abs(f.n[k]);
        }
        for (int j = 0; j < dim; ++j) den += (std::abs(v.u[j]) + c) * s[j] * 0.5;
    }
    v.dt = o.cfl * v.V / den;
}

/* ------------------------------------------------------------------------- */
/* Setup (psstGridFV::solver_parameters, psstFiniteVolumeGrid.cpp:36-194).   */
/* ------------------------------------------------------------------------- */

void solver_parameters(Oracle &o, const std::string &prefix) {
    std::ifstream in(prefix + ".inp");
    if (!in) throw std::runtime_error("cannot open " + prefix + ".inp");
    int init_mode = (int)get_parameter(in);
    initialize_field(o, prefix, init_mode);
    load_boundary(o, prefix);
    o.cfl = get_parameter(in);
    o.lc = (int)get_parameter(in);
    if (o.lc == 0) {
        /* psstUniformTime::initialize :21-43 */
        if (std::abs(o.ref_vel) < 1.0e-10) o.ref_vel = 1.0;
        if (o.dim == 2)
            o.uniform_dt = o.cfl * std::sqrt(o.min_vol) / o.ref_vel;
        else
            o.uniform_dt = o.cfl * std::pow(o.min_vol, 1.0 / 3.0) / o.ref_vel;
        for (auto &kv : o.vol_zones)
            if (kv.second.bc == 1)
                for (int vi : kv.second.items) o.vols[vi].dt = o.uniform_dt;
    } else if (o.lc != 1 && o.lc != 2) {
        throw std::runtime_error("unknown time step mode");
    }
    long mi = (long)get_parameter(in);
    o.max_iter = mi + o.iter - 1;
    o.iter_step = (long)get_parameter(in);
    o.rk_level = (int)get_parameter(in);
    int rk_mode = (int)get_parameter(in);
    if (rk_mode == 0) {
        /* psstRungeKutta::standart_coefficients, psstRungeKutta.cpp:10-49 */
        static const double t1[] = {1.0};
        static const double t2[] = {0.4242, 1.0};
        static const double t3[] = {0.1918, 0.4929, 1.0000};
        static const double t4[] = {0.1084, 0.2602, 0.5052, 1.0000};
        static const double t5[] = {0.0695, 0.1602, 0.2898, 0.5060, 1.0000};
        const double *t[] = {nullptr, t1, t2, t3, t4, t5};
        if (o.rk_level < 1 || o.rk_level > 5) throw std::runtime_error("Unknow RK order");
        o.rk.assign(t[o.rk_level], t[o.rk_level] + o.rk_level);
        std::string line;
        std::getline(in, line);
    } else {
        o.rk.assign(o.rk_level, 0.0);
        for (int i = 0; i < o.rk_level - 1; ++i) in >> o.rk[i];
        o.rk[o.rk_level - 1] = get_parameter(in);
    }
    o.flux = (int)get_parameter(in);
    o.order = (int)get_parameter(in);
    if (o.flux < 1 || o.flux > 3) throw std::runtime_error("Unknow convective flow scheme");
    if (o.order != 1 && o.order != 2) throw std::runtime_error("Unknown scheme order");
    o.grad = (int)get_parameter(in);
    if (o.grad == 1) gg_initialize(o);
    else lsq_initialize(o);
    o.limiter = (int)get_parameter(in);
    if (o.limiter < 1 || o.limiter > 4) o.limiter = 0;
    {
        // Venkatakrishnan's eps^2 = (K h)^3 with h the mean interior cell size
        double vsum = 0.0;
        long n = 0;
        for (auto &kv : o.vol_zones)
            if (kv.second.bc == 1)
                for (int vi : kv.second.items) {
                    vsum += o.vols[vi].V;
                    ++n;
                }
        const double kh = o.venkat_k * std::pow(vsum / (double)n, 1.0 / o.dim);
        o.venkat_eps2 = kh * kh * kh;
    }
}

/* ------------------------------------------------------------------------- */
/* One iteration of explicit_solve (psstFiniteVolumeGrid.cpp:279-366).        */
/* ------------------------------------------------------------------------- */

void compute_w(Oracle &o) {
    /* :279-300 */
    const int dim = o.dim;
    for (auto &kv : o.vol_zones) {
        for (int vi : kv.second.items) {
            Vol &v = o.vols[vi];
            double vsq = 0.0;
            double rho = v.rho;
            for (int j = 0; j < dim; ++j) {
                double t = v.u[j];
                v.w[j] = rho * t;
                v.R[j] = 0.0;
                vsq += t * t;
            }
            double E = o.cv * v.T + 0.5 * vsq;
            v.R[dim] = 0.0;
            v.w[dim] = rho;
            v.R[dim + 1] = 0.0;
            v.w[dim + 1] = rho * E;
        }
    }
    o.w_ready = true;
}

void update_cells(Oracle &o, int k, double *res) {
    /* :312-349 */
    const int dim = o.dim;
    const int ws = dim + 2;
    double alpha = o.rk[k];
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc != 1) continue;
        for (int vi : kv.second.items) {
            Vol &v = o.vols[vi];
            double vol = v.V;
            time_step_at(o, v);
            for (int j = 0; j < ws; ++j) {
                if (k == 0) v.w0[j] = v.w[j];
                v.w[j] = v.w0[j] - v.dt * alpha * v.R[j] / vol;
                if (res[j] < std::abs(v.R[j])) res[j] = std::abs(v.R[j]);
                v.R[j] = 0.0;
            }
            double rho = v.w[dim];
            v.rho = rho;
            double vsq = 0.0;
            for (int j = 0; j < dim; ++j) {
                v.u[j] = v.w[j] / rho;
                double vel = v.u[j];
                vsq += vel * vel;
            }
            v.T = ((v.w[dim + 1] / rho) - (vsq * 0.5)) / o.cv;
            v.p = rho * o.rm * v.T;
        }
    }
}

void apply_bcs(Oracle &o) {
    /* :351-366 */
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc == 1 || kv.second.bc == 32) continue;
        const BCDyn &d = o.dyn.at(kv.first);
        const BCThermal &t = o.thermal.at(kv.first);
        for (int gi : kv.second.items) {
            Vol &g = o.vols[gi];
            dyn_calculate_at(o, d, g);
            thermal_calculate_at(o, t, g);
            g.rho = g.p / (g.T * o.rm);
        }
    }
}

void one_iteration(Oracle &o, double *res) {
    const int ws = o.dim + 2;
    for (int j = 0; j < ws; ++j) res[j] = 0.0;
    if (o.order == 2) {
        if (o.grad == 1) gg_calculate(o);
        else lsq_calculate(o);
    }
    for (int k = 0; k < o.rk_level; ++k) {
        convective_calculate(o);
        update_cells(o, k, res);
        apply_bcs(o);
    }
}

/* Canonical state order: interior cells (bc 1 zones, ascending zone id, then
 * zone-local index), then ghosts (boundary zones, ascending id, index). */
template <class F>
void for_each_canonical(Oracle &o, F fn) {
    size_t idx = 0;
    for (auto &kv : o.vol_zones)
        if (kv.second.bc == 1)
            for (int vi : kv.second.items) fn(idx++, o.vols[vi]);
    for (auto &kv : o.vol_zones)
        if (kv.second.bc != 1 && kv.second.bc != 32)
            for (int vi : kv.second.items) fn(idx++, o.vols[vi]);
}

thread_local std::string g_last_error;

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_last_error.c_str(); }

void *oracle_open(const char *mesh_path, const char *case_prefix) {
    try {
        auto *o = new Oracle();
        RawMesh m = read_mesh(mesh_path);
        build_grid(*o, m);
        solver_parameters(*o, case_prefix);
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

/* Same as oracle_open for a mesh given as arrays (the content of a Fluent
 * file): xyz [nv][3], face_nv [nf], face_v [nf][4], face_c [nf][2],
 * face_zone [nf], cell_zone [nc], zones (file order) with id, first/last,
 * bc and '\n'-separated names. */
void *oracle_open_raw(int dim, long nv, const double *xyz, long nf, const signed char *face_nv, const long *face_v,
                      const long *face_c, const long *face_zone, long nc, const long *cell_zone, long nz,
                      const long *zone_id, const long *zone_range, const long *zone_bc, const char *zone_names,
                      const char *case_prefix) {
    try {
        RawMesh m;
        m.dim = dim;
        m.xyz.assign(xyz, xyz + 3 * nv);
        m.faces.resize((size_t)nf);
        for (long i = 0; i < nf; ++i) {
            RawFace &f = m.faces[(size_t)i];
            f.nv = face_nv[i];
            for (int k = 0; k < 4; ++k) f.v[k] = face_v[4 * i + k];
            f.c[0] = face_c[2 * i];
            f.c[1] = face_c[2 * i + 1];
            f.zone = face_zone[i];
        }
        m.cell_zone.assign(cell_zone, cell_zone + nc);
        std::string names(zone_names);
        size_t pos = 0;
        for (long z = 0; z < nz; ++z) {
            size_t e = names.find('\n', pos);
            if (e == std::string::npos) e = names.size();
            RawZone rz;
            rz.id = zone_id[z];
            rz.first = zone_range[2 * z];
            rz.last = zone_range[2 * z + 1];
            rz.bc = zone_bc[z];
            rz.name = names.substr(pos, e - pos);
            pos = e + 1;
            m.zone_pos[rz.id] = m.zones.size();
            m.zones.push_back(rz);
        }
        auto *o = new Oracle();
        build_grid(*o, m);
        solver_parameters(*o, case_prefix);
        return o;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void oracle_close(void *h) { delete static_cast<Oracle *>(h); }

int oracle_info(void *h, int *info /* [8] */) {
    Oracle &o = *static_cast<Oracle *>(h);
    int n_in = 0, n_gh = 0;
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc == 1) n_in += (int)kv.second.items.size();
        else if (kv.second.bc != 32) n_gh += (int)kv.second.items.size();
    }
    int n_faces = 0;
    for (auto &kv : o.face_zones)
        if (kv.second.bc != 31) n_faces += (int)kv.second.items.size();
    info[0] = o.dim;
    info[1] = n_in;
    info[2] = n_gh;
    info[3] = n_faces;
    info[4] = o.rk_level;
    info[5] = (int)o.iter;
    info[6] = (int)o.max_iter;
    info[7] = (int)o.iter_step;
    return 0;
}

int oracle_constants(void *h, double *c /* [8] */) {
    Oracle &o = *static_cast<Oracle *>(h);
    c[0] = o.gamma; c[1] = o.rm; c[2] = o.cp; c[3] = o.cv;
    c[4] = o.ref_vel; c[5] = o.cfl; c[6] = o.min_vol; c[7] = o.uniform_dt;
    return 0;
}

int oracle_run(void *h, int n_iter, double *res_hist) {
    Oracle &o = *static_cast<Oracle *>(h);
    try {
        if (!o.w_ready) compute_w(o);
        const int ws = o.dim + 2;
        for (int it = 0; it < n_iter; ++it) {
            one_iteration(o, res_hist ? res_hist + (size_t)it * ws : std::vector<double>(ws).data());
            ++o.iter;
        }
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return 1;
    }
}

int oracle_get_state(void *h, double *prim /* [(dim+3) * (N+G)] SoA */) {
    Oracle &o = *static_cast<Oracle *>(h);
    int info[8];
    oracle_info(h, info);
    size_t tot = (size_t)info[1] + info[2];
    const int dim = o.dim;
    for_each_canonical(o, [&](size_t i, Vol &v) {
        for (int j = 0; j < dim; ++j) prim[j * tot + i] = v.u[j];
        prim[dim * tot + i] = v.p;
        prim[(dim + 1) * tot + i] = v.T;
        prim[(dim + 2) * tot + i] = v.rho;
    });
    return 0;
}

int oracle_set_state(void *h, const double *prim) {
    Oracle &o = *static_cast<Oracle *>(h);
    int info[8];
    oracle_info(h, info);
    size_t tot = (size_t)info[1] + info[2];
    const int dim = o.dim;
    for_each_canonical(o, [&](size_t i, Vol &v) {
        for (int j = 0; j < dim; ++j) v.u[j] = prim[j * tot + i];
        v.p = prim[dim * tot + i];
        v.T = prim[(dim + 1) * tot + i];
        v.rho = prim[(dim + 2) * tot + i];
    });
    o.w_ready = false;
    return 0;
}

int oracle_get_w(void *h, double *w /* [(dim+2) * N] SoA, interior only */) {
    Oracle &o = *static_cast<Oracle *>(h);
    int info[8];
    oracle_info(h, info);
    size_t n = (size_t)info[1];
    const int ws = o.dim + 2;
    for_each_canonical(o, [&](size_t i, Vol &v) {
        if (i < n)
            for (int j = 0; j < ws; ++j) w[j * n + i] = v.w[j];
    });
    return 0;
}

/* Gradients of (u_0..u_{d-1}, p, T) of the interior cells for the current
 * state: [var][k][N] SoA. */
int oracle_gradients(void *h, double *grads) {
    Oracle &o = *static_cast<Oracle *>(h);
    try {
        if (o.grad == 1) gg_calculate(o);
        else lsq_calculate(o);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return 1;
    }
    int info[8];
    oracle_info(h, info);
    size_t n = (size_t)info[1];
    const int dim = o.dim;
    for_each_canonical(o, [&](size_t i, Vol &v) {
        if (i >= n) return;
        for (int var = 0; var < dim + 2; ++var)
            for (int k = 0; k < dim; ++k) grads[((size_t)var * dim + k) * n + i] = v.g[var][k];
    });
    return 0;
}

/* Residual R of one flux pass on the current state (gradients as last
 * computed), interior cells: [(dim+2)][N].  Leaves R zeroed afterwards. */
int oracle_residual(void *h, double *R) {
    Oracle &o = *static_cast<Oracle *>(h);
    try {
        convective_calculate(o);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return 1;
    }
    int info[8];
    oracle_info(h, info);
    size_t n = (size_t)info[1];
    const int ws = o.dim + 2;
    for_each_canonical(o, [&](size_t i, Vol &v) {
        for (int j = 0; j < ws; ++j) {
            if (i < n) R[j * n + i] = v.R[j];
            v.R[j] = 0.0;
        }
    });
    return 0;
}

/* Scalar limiter psi(num, den) (psstScalarLimiters.cpp). */
double oracle_limiter(int kind, double num, double den, double eps2) { return limiter_value(kind, num, den, eps2); }

/* Interior-face flux of two states q = (u..., p, T, rho) for unit normal n
 * pointing toward state 0; returns 0, or 1 when the reference would throw. */
int oracle_face_flux(int flux, int dim, double gamma, double rm, double cp, double cv, const double *n,
                     const double *q0, const double *q1, double *F) {
    RVar rv[2];
    const double *q[2] = {q0, q1};
    for (int j = 0; j < 2; ++j) {
        for (int k = 0; k < dim; ++k) rv[j].u[k] = q[j][k];
        rv[j].p = q[j][dim];
        rv[j].T = q[j][dim + 1];
        rv[j].rho = q[j][dim + 2];
    }
    try {
        riemann_flux(flux, dim, Gas{gamma, rm, cp, cv}, n, rv, F);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return 1;
    }
}

/* Write the reference's `.sav` format (psstFiniteVolumeGrid.cpp:606-639). */
int oracle_write_sav(void *h, const char *path, long iter_value) {
    Oracle &o = *static_cast<Oracle *>(h);
    std::ofstream out(path);
    if (!out) return 1;
    out.precision(11);
    out << iter_value << std::endl;
    for (auto &kv : o.vol_zones) {
        if (kv.second.bc != 1) continue;
        out << kv.first << std::endl;
        out << kv.second.items.size() << std::endl;
        for (int vi : kv.second.items) {
            const Vol &v = o.vols[vi];
            out << std::fixed << v.p << " ";
            out << std::fixed << v.T << " ";
            for (int j = 0; j < o.dim; ++j) out << std::fixed << v.u[j] << " ";
            out << std::endl;
        }
    }
    out << 0;
    return 0;
}

}  // extern "C"
