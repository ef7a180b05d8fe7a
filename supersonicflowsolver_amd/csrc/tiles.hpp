// tiles.hpp -- device numbering and tile plan of the fused stage kernels.
//
// Interior cells are renumbered along a Morton (Z-order) curve of their
// centroids, so that a contiguous range of device cell indices is a compact
// patch of the mesh whatever the file's numbering (column-major, RCM or a
// random permutation).  A tile is such a range; its workgroup stages the
// tile's cells and their face neighbours ("halo", interior or ghost) in LDS,
// computes every face flux of the tile once, and gathers them per cell in
// the reference's summation order.  Renumbering changes storage only: every
// per-cell sum keeps the reference's order, so FP64 results stay bitwise
// identical.
#pragma once

#include <cstdint>
#include <vector>

namespace psst {

// Device position of every interior cell (Morton order of the centroids,
// ties by canonical index).  d2c[d] = canonical index of device cell d.
// seg (optional) lists the first canonical cell of each independent
// segment (instances of a batch, seg[0] = 0): cells are ordered within their
// segment, and segments keep their order.
// axis_h (optional, dim values): the mesh's typical cell spacing per axis;
// coordinates are divided by it so that Morton boxes are cubes in cells,
// not in length (an anisotropic mesh would otherwise get slab-shaped tiles).
std::vector<int32_t> morton_order(int dim, int32_t N, int64_t NT, const double *centre /* [dim][NT] */,
                                  const std::vector<int32_t> &seg = {}, const double *axis_h = nullptr);

struct TilePlan {
    int32_t tile = 256;                 // nominal cells per tile
    int32_t n_tiles = 0;
    int32_t W_visit = 0, W_list = 0;    // ELL widths
    int32_t max_slots = 0, max_faces = 0, max_ghosts = 0;
    std::vector<int32_t> cell_begin;    // [T+1] device interior cell ranges
    std::vector<int32_t> halo_off;      // [T+1]
    std::vector<int32_t> halo;          // device ids of halo cells (interior outside the tile, or ghosts)
    std::vector<int32_t> face_off;      // [T+1]
    std::vector<int32_t> face_id;       // global face id of each tile face
    std::vector<uint32_t> face_ends;    // local slot of neighbour 0 | neighbour 1 << 15 | kind << 30
                                        // kind: 0 interior, 1 ghost is neighbour 0, 2 ghost is neighbour 1
    std::vector<int32_t> ghost_off;     // [T+1]
    std::vector<int32_t> ghost;         // ghost device ids updated by the tile
    std::vector<int16_t> visit_ell;     // [W_visit][N] local face id | 0x4000 if the cell is the face's
                                        // neighbour 0; face-loop order, -1 padded
    std::vector<int16_t> list_ell;      // [W_list][N] local face ids, face-list order, -1 padded
    std::vector<int16_t> list_slot;     // [W_list][N] local slot of the neighbour across each of them
};

// f0/f1: device cell ids per global face; voff/vface and loff/lface: per
// device interior cell, global face ids in face-loop and face-list order.
// seg (optional): first device cell of each segment; no tile spans two.
// The workgroup's LDS is sized for the largest tile, so a few irregular
// tiles (where the Morton curve jumps) would cost every workgroup: with
// cost_slot/cost_face > 0 (LDS bytes per slot and per face), tiles whose
// cost exceeds 1.15 x the median are halved until they fit.
// face_class (optional, per global face): a tile's faces are ordered by
// class (stable, so first appearance over the tile's cells within a class).
// Local face order is storage only: every per-cell sum follows the visit and
// face lists, whatever the local indices.
TilePlan plan_tiles(int32_t N, int32_t G, int32_t F, int32_t tile, const int32_t *f0, const int32_t *f1,
                    const int32_t *voff, const int32_t *vface, const int32_t *loff, const int32_t *lface,
                    const std::vector<int32_t> &seg = {}, double cost_slot = 0.0, double cost_face = 0.0,
                    const uint8_t *face_class = nullptr);

}  // namespace psst
