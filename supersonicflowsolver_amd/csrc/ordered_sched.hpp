// ordered_sched.hpp -- schedule of the reference-ordered time step (time
// step mode 1) on large meshes.
//
// The reference updates the cells in place, one after the other
// (psstFiniteVolumeGrid.cpp:317-347), and mode 1's time step of a cell
// reads the new state of the face-list neighbours that come earlier in that
// loop (psstLocalTimeCalculation.cpp:90-131).  The update is therefore a
// chain: a cell's level is 1 + the largest level of its earlier neighbours,
// and a level's cells are independent of each other.
//
// On large meshes the cells are split into chunks (compact patches).  One
// workgroup owns a chunk and walks its cells level by level.  Before a level
// it waits only for the chunks that own earlier neighbours of that level's
// cells, each to have finished a given level; a chunk publishes its progress
// only after the levels that some other chunk waits for.  A wait that an
// earlier wait of the same chunk already implies is left out.
#pragma once

#include <cstdint>
#include <vector>

namespace psst {

// Dependency level of every interior cell, canonical numbering:
// 1 + max level of its face-list neighbours j < i (ghosts, j >= N, never
// count).  lnb[k] is the neighbour across list entry k.
std::vector<int32_t> ordered_levels(int32_t N, const int32_t *list_offsets, const int32_t *lnb);

constexpr int32_t kSchedPublish = 1 << 30;  // bit of ent_level: publish progress after the entry

struct ChunkSchedule {
    int32_t n_chunks = 0;
    std::vector<int32_t> pos_cell;    // [N] device cell at each position (chunk, then level, then device id)
    std::vector<int32_t> ch_ent;      // [n_chunks+1] entries of each chunk
    std::vector<int32_t> ent_pos;     // [E+1] first position of each entry (one level of one chunk)
    std::vector<int32_t> ent_level;   // [E] level | kSchedPublish
    std::vector<int32_t> ent_wait;    // [E+1] waits before each entry
    std::vector<int32_t> wait_chunk;  // chunk waited for
    std::vector<int32_t> wait_level;  // level that chunk must have finished
};

// Partition by ranges of the device numbering (a Morton curve): chunk of
// canonical cell i = c2d[i] / chunk_cells.  Returns the chunk count.
int32_t morton_chunks(int32_t N, int32_t chunk_cells, const std::vector<int32_t> &c2d, std::vector<int32_t> &chunk_of);

// Partition into boxes of a potential per axis: phi_k(i) = max over the
// earlier neighbours j of phi_k(j) + [the face between them is most normal
// to axis k].  Every potential grows along every dependency, so a box's
// earlier neighbours lie in boxes with no larger coordinate on any axis:
// chunks never wait for each other both ways.  On a structured mesh the
// potentials are the cell's row and column, and the boxes are rectangles of
// about chunk_cells cells.  face_normal is [dim][F]; lface / lnb are the
// face-list entries.  Returns the chunk count (empty boxes dropped).
int32_t potential_chunks(int32_t N, int dim, int32_t chunk_cells, const int32_t *list_offsets,
                         const int32_t *list_faces, const int32_t *lnb, const double *face_normal, int32_t F,
                         std::vector<int32_t> &chunk_of);

// chunk_of: chunk of every interior cell (canonical); lev: canonical levels;
// c2d / d2c: canonical <-> device numbering of the interior cells.  A chunk
// publishes only after every publish_stride-th of its entries (and its
// last), the first such entry at or after a waited-for level: fewer
// publications, each a wait for the chunk's stores to complete, for a
// consumer that learns of a level up to publish_stride - 1 levels later.
ChunkSchedule chunk_schedule(int32_t N, int32_t n_chunks, const std::vector<int32_t> &chunk_of,
                             const std::vector<int32_t> &lev, const int32_t *list_offsets, const int32_t *lnb,
                             const std::vector<int32_t> &c2d, const std::vector<int32_t> &d2c,
                             int32_t publish_stride = 1);

struct ScheduleStats {
    int64_t levels = 0, chunks = 0, entries = 0, waits = 0, publishes = 0, widest_entry = 0;
    // makespan in units of one pass of a workgroup over a level (ceil(cells /
    // threads) passes per entry), with each wait across chunks costing
    // `wait_cost` more units
    int64_t makespan_free = 0, makespan_wait4 = 0;
};

// Whether every chunk finishes when the chunks run as workgroups that wait
// and publish (no cycle of waits).  With publish_stride 1 this always holds;
// a larger stride can close a cycle between chunks that wait for each other.
bool schedule_completes(const ChunkSchedule &cs);

// The figures of a schedule (no checks): counts and makespans.
ScheduleStats schedule_stats(const ChunkSchedule &cs, int threads);

// Checks a schedule against the dependency it must honour and returns its
// figures; throws std::logic_error on a violation:
//  * positions hold every interior cell once, in its own chunk; each entry's
//    cells have the entry's level; entries of a chunk rise in level;
//  * every earlier neighbour of a cell is either in the same chunk at a
//    lower level, or covered by a wait (of this or an earlier entry of the
//    chunk) for its chunk at its level or above;
//  * every wait is for a level its chunk publishes at or after;
//  * run as workgroups that wait and publish, every chunk finishes (no
//    cycle of waits).
ScheduleStats check_schedule(const ChunkSchedule &cs, int32_t N, const std::vector<int32_t> &chunk_of,
                             const std::vector<int32_t> &lev, const int32_t *list_offsets, const int32_t *lnb,
                             const std::vector<int32_t> &c2d, const std::vector<int32_t> &d2c, int threads);

}  // namespace psst
