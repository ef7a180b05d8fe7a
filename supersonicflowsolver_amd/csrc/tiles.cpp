// tiles.cpp -- Morton renumbering and the tile plan (see tiles.hpp).
#include "tiles.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace psst {

namespace {

uint64_t spread2(uint64_t x) {  // 32 bits -> every other bit of 64
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit of 63
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

}  // namespace

std::vector<int32_t> morton_order(int dim, int32_t N, int64_t NT, const double *centre,
                                  const std::vector<int32_t> &seg, const double *axis_h) {
    double inv_h[3] = {1.0, 1.0, 1.0};
    if (axis_h)
        for (int k = 0; k < dim; ++k) inv_h[k] = axis_h[k] > 0.0 ? 1.0 / axis_h[k] : 1.0;
    std::vector<int32_t> bounds = seg.empty() ? std::vector<int32_t>{0} : seg;
    bounds.push_back(N);
    std::vector<int32_t> d2c((size_t)N);
    std::vector<std::pair<uint64_t, int32_t>> key;
    for (size_t sidx = 0; sidx + 1 < bounds.size(); ++sidx) {
        const int32_t b = bounds[sidx], e = bounds[sidx + 1];
        if (b > e || b < 0 || e > N) throw std::runtime_error("segments must be ascending cell ranges");
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int k = 0; k < dim; ++k)
            for (int32_t c = b; c < e; ++c) {
                double v = centre[(size_t)k * NT + c] * inv_h[k];
                lo[k] = std::min(lo[k], v);
                hi[k] = std::max(hi[k], v);
            }
        // one common scale for all axes keeps tiles square in physical space
        double span = 0.0;
        for (int k = 0; k < dim; ++k) span = std::max(span, hi[k] - lo[k]);
        if (span <= 0.0) span = 1.0;
        const double levels = (dim == 2) ? 4294967295.0 : 2097151.0;
        key.resize((size_t)(e - b));
        for (int32_t c = b; c < e; ++c) {
            uint64_t q[3] = {0, 0, 0};
            for (int k = 0; k < dim; ++k) {
                double t = (centre[(size_t)k * NT + c] * inv_h[k] - lo[k]) / span;
                t = std::min(std::max(t, 0.0), 1.0);
                q[k] = (uint64_t)(t * levels);
            }
            uint64_t code = (dim == 2) ? (spread2(q[0]) | (spread2(q[1]) << 1))
                                       : (spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2));
            key[(size_t)(c - b)] = {code, c};
        }
        std::sort(key.begin(), key.end());
        for (int32_t d = b; d < e; ++d) d2c[(size_t)d] = key[(size_t)(d - b)].second;
    }
    return d2c;
}

namespace {

// Slots (own cells + halo) and faces of the tile [b, e).
void tile_size(int32_t b, int32_t e, const int32_t *f0, const int32_t *f1, const int32_t *voff, const int32_t *vface,
               const uint8_t *face_class, std::vector<int32_t> &slot_of, std::vector<int32_t> &lface_of,
               std::vector<int32_t> &tile_faces, std::vector<int32_t> &halo_cells) {
    const int32_t nc = e - b;
    for (int32_t c = b; c < e; ++c) slot_of[(size_t)c] = c - b;
    tile_faces.clear();
    halo_cells.clear();
    for (int32_t c = b; c < e; ++c)
        for (int32_t k = voff[c]; k < voff[c + 1]; ++k) {
            int32_t f = vface[k];
            if (lface_of[(size_t)f] < 0) {
                lface_of[(size_t)f] = (int32_t)tile_faces.size();
                tile_faces.push_back(f);
            }
        }
    if (face_class) {
        // faces of one class next to each other, so that a wavefront's faces
        // mostly take the same branches of the flux kernel
        std::stable_sort(tile_faces.begin(), tile_faces.end(),
                         [face_class](int32_t x, int32_t y) { return face_class[x] < face_class[y]; });
        for (size_t j = 0; j < tile_faces.size(); ++j) lface_of[(size_t)tile_faces[j]] = (int32_t)j;
    }
    // halo slots in order of first appearance over the tile's faces
    for (int32_t f : tile_faces)
        for (int32_t x : {f0[f], f1[f]})
            if (slot_of[(size_t)x] < 0) {
                slot_of[(size_t)x] = nc + (int32_t)halo_cells.size();
                halo_cells.push_back(x);
            }
}

void reset(int32_t b, int32_t e, std::vector<int32_t> &slot_of, std::vector<int32_t> &lface_of,
           const std::vector<int32_t> &tile_faces, const std::vector<int32_t> &halo_cells) {
    for (int32_t c = b; c < e; ++c) slot_of[(size_t)c] = -1;
    for (int32_t x : halo_cells) slot_of[(size_t)x] = -1;
    for (int32_t f : tile_faces) lface_of[(size_t)f] = -1;
}

}  // namespace

TilePlan plan_tiles(int32_t N, int32_t G, int32_t F, int32_t tile, const int32_t *f0, const int32_t *f1,
                    const int32_t *voff, const int32_t *vface, const int32_t *loff, const int32_t *lface,
                    const std::vector<int32_t> &seg, double cost_slot, double cost_face, const uint8_t *face_class) {
    TilePlan p;
    p.tile = tile;
    const int64_t NT = (int64_t)N + G;
    std::vector<int32_t> slot_of((size_t)NT, -1);   // device cell -> local slot in the current tile
    std::vector<int32_t> lface_of((size_t)F, -1);   // global face -> local face in the current tile
    std::vector<int32_t> tile_faces, halo_cells;
    // tile ranges: consecutive runs of `tile` cells, restarting at every segment
    std::vector<int32_t> starts;
    {
        std::vector<int32_t> bounds = seg.empty() ? std::vector<int32_t>{0} : seg;
        bounds.push_back(N);
        for (size_t sidx = 0; sidx + 1 < bounds.size(); ++sidx)
            for (int32_t c = bounds[sidx]; c < bounds[sidx + 1]; c += tile) starts.push_back(c);
    }
    // split the tiles whose LDS cost is well above the median
    if (cost_slot > 0.0 || cost_face > 0.0) {
        auto cost_of = [&](int32_t b, int32_t e) {
            tile_size(b, e, f0, f1, voff, vface, face_class, slot_of, lface_of, tile_faces, halo_cells);
            const double c = cost_slot * (double)((e - b) + (int32_t)halo_cells.size()) + cost_face * (double)tile_faces.size();
            reset(b, e, slot_of, lface_of, tile_faces, halo_cells);
            return c;
        };
        const size_t T0 = starts.size();
        std::vector<double> cost(T0);
        for (size_t t = 0; t < T0; ++t) cost[t] = cost_of(starts[t], t + 1 < T0 ? starts[t + 1] : N);
        std::vector<double> sorted = cost;
        std::nth_element(sorted.begin(), sorted.begin() + sorted.size() / 2, sorted.end());
        const double cap = 1.15 * sorted[sorted.size() / 2];
        std::vector<int32_t> refined;
        refined.reserve(T0);
        for (size_t t = 0; t < T0; ++t) {
            const int32_t b = starts[t], e = t + 1 < T0 ? starts[t + 1] : N;
            if (cost[t] <= cap) {
                refined.push_back(b);
                continue;
            }
            // halve until every piece fits (or is 32 cells)
            std::vector<std::pair<int32_t, int32_t>> stack{{b, e}};
            std::vector<int32_t> pieces;
            while (!stack.empty()) {
                auto r = stack.back();
                stack.pop_back();
                if (r.second - r.first > 32 && cost_of(r.first, r.second) > cap) {
                    const int32_t mid = r.first + (r.second - r.first) / 2;
                    stack.push_back({mid, r.second});
                    stack.push_back({r.first, mid});
                } else {
                    pieces.push_back(r.first);
                }
            }
            refined.insert(refined.end(), pieces.begin(), pieces.end());
        }
        starts.swap(refined);
    }
    p.n_tiles = (int32_t)starts.size();
    for (int32_t c = 0; c < N; ++c) {
        p.W_visit = std::max(p.W_visit, voff[c + 1] - voff[c]);
        p.W_list = std::max(p.W_list, loff[c + 1] - loff[c]);
    }
    p.visit_ell.assign((size_t)p.W_visit * N, (int16_t)-1);
    p.list_ell.assign((size_t)p.W_list * N, (int16_t)-1);
    p.list_slot.assign((size_t)p.W_list * N, (int16_t)-1);
    p.cell_begin.resize((size_t)p.n_tiles + 1);
    p.halo_off.assign(1, 0);
    p.face_off.assign(1, 0);
    p.ghost_off.assign(1, 0);
    for (int32_t t = 0; t < p.n_tiles; ++t) {
        const int32_t b = starts[(size_t)t], e = (t + 1 < p.n_tiles) ? starts[(size_t)t + 1] : N;
        p.cell_begin[(size_t)t] = b;
        const int32_t nc = e - b;
        tile_size(b, e, f0, f1, voff, vface, face_class, slot_of, lface_of, tile_faces, halo_cells);
        const int32_t ns = nc + (int32_t)halo_cells.size();
        const int32_t nf = (int32_t)tile_faces.size();
        if (ns >= 32768 || nf >= 16384) throw std::runtime_error("tile too large for its local indices");
        p.max_slots = std::max(p.max_slots, ns);
        p.max_faces = std::max(p.max_faces, nf);
        for (int32_t f : tile_faces) {
            uint32_t l0 = (uint32_t)slot_of[(size_t)f0[f]], l1 = (uint32_t)slot_of[(size_t)f1[f]];
            uint32_t kind = (f0[f] >= N) ? 1u : (f1[f] >= N ? 2u : 0u);
            p.face_ends.push_back(l0 | (l1 << 15) | (kind << 30));
            p.face_id.push_back(f);
            if (kind != 0) {  // the tile updates this ghost after its cells
                int32_t gcell = (kind == 1) ? f0[f] : f1[f];
                p.ghost.push_back(gcell);
            }
        }
        int32_t ng = (int32_t)(p.ghost.size() - (size_t)p.ghost_off.back());
        p.max_ghosts = std::max(p.max_ghosts, ng);
        for (int32_t c = b; c < e; ++c) {
            int32_t w = 0;
            for (int32_t k = voff[c]; k < voff[c + 1]; ++k, ++w) {
                const int32_t f = vface[k];
                p.visit_ell[(size_t)w * N + c] = (int16_t)(lface_of[(size_t)f] | (f0[f] == c ? 0x4000 : 0));
            }
            w = 0;
            for (int32_t k = loff[c]; k < loff[c + 1]; ++k, ++w) {
                const int32_t f = lface[k];
                int32_t lf = lface_of[(size_t)f];
                if (lf < 0) throw std::runtime_error("face list entry outside the tile's faces");
                p.list_ell[(size_t)w * N + c] = (int16_t)lf;
                p.list_slot[(size_t)w * N + c] = (int16_t)slot_of[(size_t)(f0[f] == c ? f1[f] : f0[f])];
            }
        }
        p.halo.insert(p.halo.end(), halo_cells.begin(), halo_cells.end());
        p.halo_off.push_back((int32_t)p.halo.size());
        p.face_off.push_back((int32_t)p.face_id.size());
        p.ghost_off.push_back((int32_t)p.ghost.size());
        reset(b, e, slot_of, lface_of, tile_faces, halo_cells);
    }
    p.cell_begin[(size_t)p.n_tiles] = N;
    return p;
}

}  // namespace psst
