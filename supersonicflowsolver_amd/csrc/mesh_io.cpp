// mesh_io.cpp -- Fluent ASCII mesh reader and writer.
//
// Semantics follow the reference reader (psstAnsysDirector::build_step and
// psstAnsysBuilder::read_*, psstAnsysGridReader.cpp:948-1116, :1138-1306):
//   * section ids are decimal, header fields and indices hexadecimal,
//     vertex coordinates decimal;
//   * a face line is [type if the zone is mixed] v0 v1 [v2 v3] c0 c1 and the
//     first cell index becomes cell(0), the second cell(1)
//     (read_faces :1019-1026 with psstSimpleFace's (left, right) swap :177-194);
//   * zones are kept in order of first appearance (add_zones_storage :600-632),
//     a tree section counting as an appearance of its parent and child zones
//     (add_zones_storage_connection :655-686).
// Unlike the reference it parses by structure rather than by lines, so it is
// indifferent to CRLF (quirk Q12) and to where the opening '(' of a data
// block sits.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>

#include "psst_host.hpp"

namespace psst {

const RawZone *RawMesh::zone(int64_t id) const {
    for (const RawZone &z : zones)
        if (z.id == id) return &z;
    return nullptr;
}

namespace {

struct Scanner {
    const char *p;
    const char *end;
    void ws() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool peek(char c) {
        ws();
        return p < end && *p == c;
    }
    void expect(char c) {
        ws();
        if (p >= end || *p != c) throw std::runtime_error(std::string("mesh: expected '") + c + "'");
        ++p;
    }
    bool is_delim(char c) const { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '(' || c == ')'; }
    int64_t hex() {
        ws();
        int64_t v = 0;
        const char *b = p;
        while (p < end) {
            char c = *p;
            int d;
            if (c >= '0' && c <= '9') d = c - '0';
            else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
            else break;
            v = v * 16 + d;
            ++p;
        }
        if (p == b) throw std::runtime_error("mesh: expected a hexadecimal integer");
        return v;
    }
    int64_t dec() {
        ws();
        int64_t v = 0;
        const char *b = p;
        while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
        if (p == b) throw std::runtime_error("mesh: expected an integer");
        return v;
    }
    double real() {
        ws();
        char *e = nullptr;
        double v = std::strtod(p, &e);
        if (e == p) throw std::runtime_error("mesh: expected a real number");
        p = e;
        return v;
    }
    std::string word() {
        ws();
        const char *b = p;
        while (p < end && !is_delim(*p)) ++p;
        return std::string(b, p);
    }
    std::vector<int64_t> header() {
        expect('(');
        std::vector<int64_t> v;
        while (!peek(')')) v.push_back(hex());
        ++p;
        return v;
    }
    void skip_close() {  // skip to the ')' that closes the current level
        int depth = 1;
        while (p < end && depth > 0) {
            if (*p == '"') {
                ++p;
                while (p < end && *p != '"') ++p;
            } else if (*p == '(') {
                ++depth;
            } else if (*p == ')') {
                --depth;
            }
            ++p;
        }
    }
};

void add_zone(RawMesh &m, std::map<int64_t, size_t> &pos, int64_t id, int64_t first, int64_t last, int64_t bc) {
    auto it = pos.find(id);
    if (it == pos.end()) {
        RawZone z;
        z.id = id;
        z.first = first;
        z.last = last;
        z.bc = bc;
        pos[id] = m.zones.size();
        m.zones.push_back(z);
    } else {
        RawZone &z = m.zones[it->second];
        z.first = first;
        z.last = last;
        z.bc = bc;
    }
}

// add_zones_storage_connection (psstAnsysGridReader.cpp:655-686): a parent
// zone first seen in a tree section is created as a parent face zone (type
// 31) with the tree's range, a child zone first seen there as an empty zone;
// both keep that place in the zone order when their own section arrives.
void add_zone_link(RawMesh &m, std::map<int64_t, size_t> &pos, int64_t parent, int64_t first, int64_t last,
                   int64_t child) {
    auto it = pos.find(parent);
    if (it == pos.end()) {
        RawZone z;
        z.id = parent;
        z.first = first;
        z.last = last;
        z.bc = 31;
        z.child_zone = child;
        pos[parent] = m.zones.size();
        m.zones.push_back(z);
    } else {
        m.zones[it->second].child_zone = child;
    }
    it = pos.find(child);
    if (it == pos.end()) {
        RawZone z;
        z.parent_zone = parent;
        pos[child] = m.zones.size();
        m.zones.push_back(z);
        m.zones.back().id = child;
    } else {
        m.zones[it->second].parent_zone = parent;
    }
}

}  // namespace

RawMesh read_fluent_msh(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("Error: Could not open \"" + path + "\" file!");
    std::string text;
    text.resize((size_t)f.tellg());
    f.seekg(0);
    f.read(&text[0], (std::streamsize)text.size());
    RawMesh m;
    std::map<int64_t, size_t> pos;
    Scanner s{text.data(), text.data() + text.size()};
    int nd_hint = 0;
    while (true) {
        s.ws();
        if (s.p >= s.end) break;
        if (*s.p != '(') {
            ++s.p;
            continue;
        }
        ++s.p;
        int64_t sec = s.dec();
        if (sec == 2) {
            m.dim = (int)s.dec();
            s.skip_close();
        } else if (sec == 10) {
            std::vector<int64_t> h = s.header();
            if (h.size() < 3) throw std::runtime_error("mesh: bad node section");
            if (h[0] == 0) {
                if (h.size() > 4) nd_hint = (int)h[4];
                s.skip_close();
                continue;
            }
            int dim = m.dim ? m.dim : (h.size() > 4 ? (int)h[4] : nd_hint);
            if (dim != 2 && dim != 3) throw std::runtime_error("mesh: node section before dimension");
            if (m.n_vertices() != h[1] - 1) throw std::runtime_error("mesh: node sections out of order");
            int64_t n = h[2] - h[1] + 1;
            s.expect('(');
            m.xyz.reserve(m.xyz.size() + 3 * (size_t)n);
            for (int64_t i = 0; i < n; ++i) {
                double x[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < dim; ++k) x[k] = s.real();
                m.xyz.insert(m.xyz.end(), x, x + 3);
            }
            s.expect(')');
            s.skip_close();
        } else if (sec == 12) {
            std::vector<int64_t> h = s.header();
            if (h.size() < 3) throw std::runtime_error("mesh: bad cell section");
            if (h[0] == 0) {
                s.skip_close();
                continue;
            }
            if (h.size() < 5) throw std::runtime_error("mesh: bad cell section");
            add_zone(m, pos, h[0], h[1], h[2], h[3]);
            if (m.n_cells() < h[2]) {
                m.cell_zone.resize((size_t)h[2], 0);
                m.cell_type.resize((size_t)h[2], 0);
            }
            for (int64_t i = h[1]; i <= h[2]; ++i) {
                m.cell_zone[(size_t)i - 1] = h[0];
                m.cell_type[(size_t)i - 1] = (int8_t)h[4];
            }
            if (s.peek('(')) {
                ++s.p;
                if (h[4] == 0)
                    for (int64_t i = h[1]; i <= h[2]; ++i) m.cell_type[(size_t)i - 1] = (int8_t)s.hex();
                s.expect(')');
            }
            s.skip_close();
        } else if (sec == 13) {
            std::vector<int64_t> h = s.header();
            if (h.size() < 3) throw std::runtime_error("mesh: bad face section");
            if (h[0] == 0) {
                s.skip_close();
                continue;
            }
            if (h.size() < 5) throw std::runtime_error("mesh: bad face section");
            add_zone(m, pos, h[0], h[1], h[2], h[3]);
            // Faces go to their declared indices, so zone sections may come in
            // any order (the reference's own .cas writer emits them by zone
            // id); the reference appends them in reading order instead.
            if (h[1] < 1 || h[2] < h[1]) throw std::runtime_error("mesh: bad face section");
            if (m.n_faces() < h[2]) {
                m.face_nv.resize((size_t)h[2], 0);
                m.face_v.resize(4 * (size_t)h[2], 0);
                m.face_c.resize(2 * (size_t)h[2], 0);
                m.face_zone.resize((size_t)h[2], 0);
            }
            s.expect('(');
            for (int64_t i = h[1] - 1; i < h[2]; ++i) {
                int64_t t = h[4];
                if (t == 0) t = s.hex();
                if (t < 2 || t > 4) throw std::runtime_error("Error: Unknown face type!");
                if (m.face_nv[(size_t)i] != 0) throw std::runtime_error("mesh: face listed twice");
                int64_t *v = &m.face_v[4 * (size_t)i];
                for (int k = 0; k < (int)t; ++k) v[k] = s.hex();
                m.face_c[2 * (size_t)i] = s.hex();
                m.face_c[2 * (size_t)i + 1] = s.hex();
                m.face_nv[(size_t)i] = (int8_t)t;
                m.face_zone[(size_t)i] = h[0];
            }
            s.expect(')');
            s.skip_close();
        } else if (sec == 39 || sec == 45) {
            s.expect('(');
            int64_t id = s.dec();
            std::string kind = s.word();
            std::string name = s.word();
            s.skip_close();
            s.skip_close();
            auto it = pos.find(id);
            if (it == pos.end()) {
                RawZone z;
                z.id = id;
                pos[id] = m.zones.size();
                m.zones.push_back(z);
                it = pos.find(id);
            }
            m.zones[it->second].kind = kind;
            m.zones[it->second].name = name;
        } else if (sec == 58 || sec == 59) {
            // (58|59 (first last parent-zone child-zone)( n kid... per parent ))
            // (read_cell_tree / read_face_tree, psstAnsysGridReader.cpp:1068-1116)
            std::vector<int64_t> h = s.header();
            if (h.size() < 4 || h[0] < 1 || h[1] < h[0]) throw std::runtime_error("mesh: bad tree section");
            add_zone_link(m, pos, h[2], h[0], h[1], h[3]);
            std::vector<std::vector<int64_t>> &kids = (sec == 58) ? m.cell_kids : m.face_kids;
            if (kids.size() < (size_t)h[1]) kids.resize((size_t)h[1]);
            s.expect('(');
            for (int64_t i = h[0]; i <= h[1]; ++i) {
                int64_t n = s.hex();
                if (n < 0 || n > (1 << 20)) throw std::runtime_error("mesh: bad tree section");
                for (int64_t j = 0; j < n; ++j) kids[(size_t)i - 1].push_back(s.hex());
            }
            s.expect(')');
            s.skip_close();
        } else {
            s.skip_close();
        }
    }
    if (m.dim == 0) m.dim = nd_hint;
    if (m.dim != 2 && m.dim != 3) throw std::runtime_error("mesh: dimension missing");
    for (int64_t z : m.cell_zone)
        if (z == 0) throw std::runtime_error("mesh: cell without zone");
    for (int8_t nv : m.face_nv)
        if (nv == 0) throw std::runtime_error("mesh: face index without a face");
    return m;
}

void write_fluent_msh(const RawMesh &m, const std::string &path) {
    for (const RawZone &z : m.zones)
        if (z.bc == 31 || z.bc == 32) throw std::runtime_error("mesh writer: adapted meshes are not supported");
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::vector<char> buf(1 << 22);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    std::fprintf(f, "(0 \"Grid\")\n\n(0 \"Dimension:\")\n(2 %d)\n\n", m.dim);
    // cells
    std::fprintf(f, "(0 \"Cells:\")\n(12 (0 1 %llx 0))\n", (unsigned long long)m.n_cells());
    for (const RawZone &z : m.zones) {
        if (z.bc != 1) continue;
        // one element type per zone, no per-cell list; the blank line after
        // the header is what the reference reader consumes for a header that
        // does not end with '(' (psstAnsysGridReader.cpp:1191-1203)
        int etype = (m.dim == 2) ? 3 : (m.face_nv.empty() || m.face_nv[0] == 4 ? 4 : 2);
        if (!m.cell_type.empty()) etype = m.cell_type[(size_t)z.first - 1];
        std::fprintf(f, "(12 (%llx %llx %llx 1 %x))\n\n", (unsigned long long)z.id, (unsigned long long)z.first,
                     (unsigned long long)z.last, etype);
    }
    // faces
    std::fprintf(f, "\n(0 \"Faces:\")\n(13 (0 1 %llx 0))\n", (unsigned long long)m.n_faces());
    for (const RawZone &z : m.zones) {
        if (z.bc == 1) continue;
        if (z.first == 0) continue;
        // all faces of a zone share one vertex count when possible
        int nv0 = m.face_nv[(size_t)z.first - 1];
        bool uniform = true;
        for (int64_t i = z.first; i <= z.last; ++i)
            if (m.face_nv[(size_t)i - 1] != nv0) uniform = false;
        std::fprintf(f, "(13 (%llx %llx %llx %llx %x)(\n", (unsigned long long)z.id, (unsigned long long)z.first,
                     (unsigned long long)z.last, (unsigned long long)z.bc, uniform ? nv0 : 0);
        for (int64_t i = z.first; i <= z.last; ++i) {
            size_t fi = (size_t)i - 1;
            int nv = m.face_nv[fi];
            if (!uniform) std::fprintf(f, " %x", nv);
            for (int k = 0; k < nv; ++k) std::fprintf(f, " %llx", (unsigned long long)m.face_v[4 * fi + k]);
            std::fprintf(f, " %llx %llx\n", (unsigned long long)m.face_c[2 * fi], (unsigned long long)m.face_c[2 * fi + 1]);
        }
        std::fputs(")\n)\n", f);
    }
    // nodes
    std::fprintf(f, "\n(0 \"Nodes:\")\n(10 (0 1 %llx 1 %d))\n", (unsigned long long)m.n_vertices(), m.dim);
    std::fprintf(f, "(10 (1 1 %llx 1 %d)(\n", (unsigned long long)m.n_vertices(), m.dim);
    for (int64_t i = 0; i < m.n_vertices(); ++i) {
        if (m.dim == 2)
            std::fprintf(f, " %.20e  %.20e\n", m.xyz[3 * i], m.xyz[3 * i + 1]);
        else
            std::fprintf(f, " %.20e  %.20e  %.20e\n", m.xyz[3 * i], m.xyz[3 * i + 1], m.xyz[3 * i + 2]);
    }
    std::fputs(")\n)\n", f);
    // zones
    std::fputs("\n(0 \"Zones:\")\n", f);
    for (const RawZone &z : m.zones) {
        std::fprintf(f, "(39 (%lld %s %s)())\n", (long long)z.id, z.kind.empty() ? "wall" : z.kind.c_str(),
                     z.name.c_str());
    }
    if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + path);
}

}  // namespace psst
