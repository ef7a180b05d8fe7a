// ordered_sched.cpp -- levels, chunk partitions and chunk schedule of the
// reference-ordered time step (see ordered_sched.hpp).
#include "ordered_sched.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>

namespace psst {

std::vector<int32_t> ordered_levels(int32_t N, const int32_t *list_offsets, const int32_t *lnb) {
    std::vector<int32_t> lev((size_t)N, 0);
    for (int32_t i = 0; i < N; ++i) {
        int32_t l = 0;
        for (int32_t k = list_offsets[i]; k < list_offsets[i + 1]; ++k) {
            const int32_t nb = lnb[k];
            if (nb < i) l = std::max(l, lev[(size_t)nb] + 1);
        }
        lev[(size_t)i] = l;
    }
    return lev;
}

int32_t morton_chunks(int32_t N, int32_t chunk_cells, const std::vector<int32_t> &c2d, std::vector<int32_t> &chunk_of) {
    if (chunk_cells <= 0) throw std::invalid_argument("chunk_cells must be positive");
    chunk_of.resize((size_t)N);
    for (int32_t i = 0; i < N; ++i) chunk_of[(size_t)i] = c2d[(size_t)i] / chunk_cells;
    return (N + chunk_cells - 1) / chunk_cells;
}

int32_t potential_chunks(int32_t N, int dim, int32_t chunk_cells, const int32_t *list_offsets,
                         const int32_t *list_faces, const int32_t *lnb, const double *face_normal, int32_t F,
                         std::vector<int32_t> &chunk_of) {
    if (chunk_cells <= 0) throw std::invalid_argument("chunk_cells must be positive");
    std::vector<int32_t> phi((size_t)dim * N, 0);
    int32_t range[3] = {1, 1, 1};
    for (int32_t i = 0; i < N; ++i) {
        int32_t v[3] = {0, 0, 0};
        for (int32_t k = list_offsets[i]; k < list_offsets[i + 1]; ++k) {
            const int32_t nb = lnb[k];
            if (nb >= i) continue;
            const int32_t f = list_faces[k];
            int ax = 0;
            for (int a = 1; a < dim; ++a)
                if (std::abs(face_normal[(size_t)a * F + f]) > std::abs(face_normal[(size_t)ax * F + f])) ax = a;
            for (int a = 0; a < dim; ++a) v[a] = std::max(v[a], phi[(size_t)a * N + nb] + (a == ax ? 1 : 0));
        }
        for (int a = 0; a < dim; ++a) {
            phi[(size_t)a * N + i] = v[a];
            range[a] = std::max(range[a], v[a] + 1);
        }
    }
    const int32_t B = std::max<int32_t>(1, (int32_t)std::lround(std::pow((double)chunk_cells, 1.0 / dim)));
    int64_t nb_box[3] = {1, 1, 1};
    for (int a = 0; a < dim; ++a) nb_box[a] = (range[a] + B - 1) / B;
    std::vector<int64_t> box((size_t)N);
    for (int32_t i = 0; i < N; ++i) {
        int64_t id = 0;
        for (int a = 0; a < dim; ++a) id = id * nb_box[a] + phi[(size_t)a * N + i] / B;
        box[(size_t)i] = id;
    }
    std::vector<int64_t> ids(box);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    chunk_of.resize((size_t)N);
    for (int32_t i = 0; i < N; ++i)
        chunk_of[(size_t)i] = (int32_t)(std::lower_bound(ids.begin(), ids.end(), box[(size_t)i]) - ids.begin());
    return (int32_t)ids.size();
}

ChunkSchedule chunk_schedule(int32_t N, int32_t n_chunks, const std::vector<int32_t> &chunk_of,
                             const std::vector<int32_t> &lev, const int32_t *list_offsets, const int32_t *lnb,
                             const std::vector<int32_t> &c2d, const std::vector<int32_t> &d2c,
                             int32_t publish_stride) {
    if (publish_stride < 1) throw std::invalid_argument("publish_stride must be positive");
    ChunkSchedule cs;
    cs.n_chunks = n_chunks;
    // cells of each chunk in device order (counting sort by chunk)
    std::vector<int32_t> start((size_t)n_chunks + 1, 0), by_chunk((size_t)N);
    for (int32_t i = 0; i < N; ++i) {
        const int32_t w = chunk_of[(size_t)i];
        if (w < 0 || w >= n_chunks) throw std::invalid_argument("chunk of a cell out of range");
        start[(size_t)w + 1]++;
    }
    for (int32_t w = 0; w < n_chunks; ++w) start[(size_t)w + 1] += start[(size_t)w];
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int32_t d = 0; d < N; ++d) by_chunk[(size_t)fill[(size_t)chunk_of[(size_t)d2c[(size_t)d]]]++] = d;
    }
    // positions: chunk by chunk, then by level (counting sort), then by device id
    cs.pos_cell.resize((size_t)N);
    cs.ch_ent.assign((size_t)n_chunks + 1, 0);
    std::vector<int32_t> count;
    for (int32_t w = 0; w < n_chunks; ++w) {
        const int32_t b = start[(size_t)w], e = start[(size_t)w + 1];
        int32_t lo = INT32_MAX, hi = -1;
        for (int32_t q = b; q < e; ++q) {
            const int32_t l = lev[(size_t)d2c[(size_t)by_chunk[(size_t)q]]];
            lo = std::min(lo, l);
            hi = std::max(hi, l);
        }
        if (e > b) {
            count.assign((size_t)(hi - lo + 2), 0);
            for (int32_t q = b; q < e; ++q) count[(size_t)(lev[(size_t)d2c[(size_t)by_chunk[(size_t)q]]] - lo + 1)]++;
            for (int32_t l = 0; l <= hi - lo; ++l) {
                if (count[(size_t)l + 1] > 0) {
                    cs.ent_pos.push_back(b + count[(size_t)l]);
                    cs.ent_level.push_back(lo + l);
                }
                count[(size_t)l + 1] += count[(size_t)l];
            }
            for (int32_t q = b; q < e; ++q) {
                const int32_t d = by_chunk[(size_t)q];
                cs.pos_cell[(size_t)(b + count[(size_t)(lev[(size_t)d2c[(size_t)d]] - lo)]++)] = d;
            }
        }
        cs.ch_ent[(size_t)w + 1] = (int32_t)cs.ent_level.size();
    }
    cs.ent_pos.push_back(N);
    // waits: per entry, for each other chunk owning an earlier neighbour of
    // one of its cells, the largest level among those neighbours; left out
    // when an earlier entry of the chunk already waited for that level or more
    const int32_t E = (int32_t)cs.ent_level.size();
    cs.ent_wait.assign((size_t)E + 1, 0);
    std::vector<std::pair<int32_t, int32_t>> need, waited;  // (chunk, level)
    for (int32_t w = 0; w < n_chunks; ++w) {
        waited.clear();
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            need.clear();
            for (int32_t p = cs.ent_pos[(size_t)en]; p < cs.ent_pos[(size_t)en + 1]; ++p) {
                const int32_t i = d2c[(size_t)cs.pos_cell[(size_t)p]];
                for (int32_t k = list_offsets[i]; k < list_offsets[i + 1]; ++k) {
                    const int32_t nb = lnb[k];
                    if (nb >= i) continue;
                    const int32_t pc = chunk_of[(size_t)nb];
                    if (pc == w) continue;
                    auto it = std::find_if(need.begin(), need.end(), [&](const auto &x) { return x.first == pc; });
                    if (it == need.end()) need.push_back({pc, lev[(size_t)nb]});
                    else it->second = std::max(it->second, lev[(size_t)nb]);
                }
            }
            std::sort(need.begin(), need.end());
            for (const auto &nd : need) {
                auto it = std::find_if(waited.begin(), waited.end(), [&](const auto &x) { return x.first == nd.first; });
                if (it != waited.end() && it->second >= nd.second) continue;
                if (it == waited.end()) waited.push_back(nd);
                else it->second = nd.second;
                cs.wait_chunk.push_back(nd.first);
                cs.wait_level.push_back(nd.second);
            }
            cs.ent_wait[(size_t)en + 1] = (int32_t)cs.wait_chunk.size();
        }
    }
    // a waited-for level publishes at the chunk's next publication point
    for (size_t i = 0; i < cs.wait_chunk.size(); ++i) {
        const int32_t pc = cs.wait_chunk[i], L = cs.wait_level[i];
        const int32_t eb = cs.ch_ent[(size_t)pc], ee = cs.ch_ent[(size_t)pc + 1];
        auto b = cs.ent_level.begin() + eb, e = cs.ent_level.begin() + ee;
        auto it = std::lower_bound(b, e, L, [](int32_t a, int32_t l) { return (a & (kSchedPublish - 1)) < l; });
        if (it == e || (*it & (kSchedPublish - 1)) != L) throw std::logic_error("chunk schedule: waited level missing");
        int32_t en = (int32_t)(it - cs.ent_level.begin());
        while (en + 1 < ee && (en - eb + 1) % publish_stride != 0) ++en;
        cs.ent_level[(size_t)en] |= kSchedPublish;
    }
    (void)c2d;
    return cs;
}

namespace {

void fail(const std::string &what) { throw std::logic_error("chunk schedule: " + what); }

constexpr int32_t kLevelMask = kSchedPublish - 1;

// The entry of chunk pc that publishes a level >= L (-1: none).
int32_t publish_entry(const ChunkSchedule &cs, int32_t pc, int32_t L) {
    for (int32_t en = cs.ch_ent[(size_t)pc]; en < cs.ch_ent[(size_t)pc + 1]; ++en)
        if ((cs.ent_level[(size_t)en] & kSchedPublish) && (cs.ent_level[(size_t)en] & kLevelMask) >= L) return en;
    return -1;
}

}  // namespace

ScheduleStats schedule_stats(const ChunkSchedule &cs, int threads) {
    const int32_t C = cs.n_chunks, E = (int32_t)cs.ent_level.size();
    ScheduleStats st;
    st.chunks = C;
    st.entries = E;
    st.waits = (int64_t)cs.wait_chunk.size();
    std::vector<int32_t> chunk_of_entry((size_t)E);
    int32_t maxlev = -1;
    for (int32_t w = 0; w < C; ++w)
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            chunk_of_entry[(size_t)en] = w;
            maxlev = std::max(maxlev, cs.ent_level[(size_t)en] & kLevelMask);
            if (cs.ent_level[(size_t)en] & kSchedPublish) st.publishes++;
            st.widest_entry = std::max<int64_t>(st.widest_entry, cs.ent_pos[(size_t)en + 1] - cs.ent_pos[(size_t)en]);
        }
    st.levels = maxlev + 1;
    std::vector<int32_t> wait_src(cs.wait_chunk.size());
    for (size_t i = 0; i < cs.wait_chunk.size(); ++i) wait_src[i] = publish_entry(cs, cs.wait_chunk[i], cs.wait_level[i]);
    // makespan: entries in level order (a wait is always for a lower level)
    std::vector<int32_t> order((size_t)E);
    for (int32_t e = 0; e < E; ++e) order[(size_t)e] = e;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return (cs.ent_level[(size_t)a] & kLevelMask) < (cs.ent_level[(size_t)b] & kLevelMask);
    });
    for (int wait_cost : {0, 4}) {
        std::vector<int64_t> fin((size_t)E, 0);
        int64_t span = 0;
        for (int32_t en : order) {
            const int32_t w = chunk_of_entry[(size_t)en];
            int64_t start = en > cs.ch_ent[(size_t)w] ? fin[(size_t)en - 1] : 0;
            for (int32_t i = cs.ent_wait[(size_t)en]; i < cs.ent_wait[(size_t)en + 1]; ++i)
                if (wait_src[(size_t)i] >= 0) start = std::max(start, fin[(size_t)wait_src[(size_t)i]] + wait_cost);
            const int64_t cells = cs.ent_pos[(size_t)en + 1] - cs.ent_pos[(size_t)en];
            fin[(size_t)en] = start + (cells + threads - 1) / threads;
            span = std::max(span, fin[(size_t)en]);
        }
        (wait_cost == 0 ? st.makespan_free : st.makespan_wait4) = span;
    }
    return st;
}

ScheduleStats check_schedule(const ChunkSchedule &cs, int32_t N, const std::vector<int32_t> &chunk_of,
                             const std::vector<int32_t> &lev, const int32_t *list_offsets, const int32_t *lnb,
                             const std::vector<int32_t> &c2d, const std::vector<int32_t> &d2c, int threads) {
    const int32_t C = cs.n_chunks;
    const int32_t E = (int32_t)cs.ent_level.size();
    if (C <= 0 || (int32_t)chunk_of.size() != N) fail("chunk count");
    if ((int32_t)cs.ch_ent.size() != C + 1 || cs.ch_ent[0] != 0 || cs.ch_ent[(size_t)C] != E) fail("ch_ent");
    if ((int32_t)cs.ent_pos.size() != E + 1 || cs.ent_pos[0] != 0 || cs.ent_pos[(size_t)E] != N) fail("ent_pos");
    if ((int32_t)cs.ent_wait.size() != E + 1 || cs.ent_wait[0] != 0 ||
        cs.ent_wait[(size_t)E] != (int32_t)cs.wait_chunk.size() || cs.wait_level.size() != cs.wait_chunk.size())
        fail("ent_wait");
    if ((int32_t)cs.pos_cell.size() != N) fail("pos_cell size");
    std::vector<char> seen((size_t)N, 0);
    for (int32_t w = 0; w < C; ++w) {
        if (cs.ch_ent[(size_t)w + 1] < cs.ch_ent[(size_t)w]) fail("ch_ent not monotone");
        int32_t prev = -1;
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            const int32_t l = cs.ent_level[(size_t)en] & kLevelMask;
            if (l <= prev) fail("levels of a chunk not rising");
            prev = l;
            if (cs.ent_pos[(size_t)en + 1] <= cs.ent_pos[(size_t)en]) fail("empty entry");
            if (cs.ent_wait[(size_t)en + 1] < cs.ent_wait[(size_t)en]) fail("ent_wait not monotone");
            for (int32_t p = cs.ent_pos[(size_t)en]; p < cs.ent_pos[(size_t)en + 1]; ++p) {
                const int32_t d = cs.pos_cell[(size_t)p];
                if (d < 0 || d >= N || seen[(size_t)d]) fail("position table");
                seen[(size_t)d] = 1;
                if (chunk_of[(size_t)d2c[(size_t)d]] != w) fail("cell outside its chunk");
                if (lev[(size_t)d2c[(size_t)d]] != l) fail("cell level differs from its entry");
            }
        }
    }
    // every wait is for a level its chunk publishes at or after
    for (size_t i = 0; i < cs.wait_chunk.size(); ++i) {
        const int32_t pc = cs.wait_chunk[i];
        if (pc < 0 || pc >= C) fail("waited chunk out of range");
        if (publish_entry(cs, pc, cs.wait_level[i]) < 0) fail("wait for a level that is never published");
    }
    // every earlier neighbour is in an earlier entry of the chunk or covered
    // by a wait of this or an earlier entry
    std::vector<int32_t> covered((size_t)C, -1);
    for (int32_t w = 0; w < C; ++w) {
        std::vector<int32_t> touched;
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            for (int32_t i = cs.ent_wait[(size_t)en]; i < cs.ent_wait[(size_t)en + 1]; ++i) {
                const int32_t pc = cs.wait_chunk[(size_t)i];
                if (pc == w) fail("chunk waits for itself");
                covered[(size_t)pc] = std::max(covered[(size_t)pc], cs.wait_level[(size_t)i]);
                touched.push_back(pc);
            }
            const int32_t l = cs.ent_level[(size_t)en] & kLevelMask;
            for (int32_t p = cs.ent_pos[(size_t)en]; p < cs.ent_pos[(size_t)en + 1]; ++p) {
                const int32_t i = d2c[(size_t)cs.pos_cell[(size_t)p]];
                for (int32_t k = list_offsets[i]; k < list_offsets[i + 1]; ++k) {
                    const int32_t nb = lnb[k];
                    if (nb >= i) continue;
                    if (lev[(size_t)nb] >= l) fail("earlier neighbour not at a lower level");
                    const int32_t pc = chunk_of[(size_t)nb];
                    if (pc != w && covered[(size_t)pc] < lev[(size_t)nb])
                        fail("cell " + std::to_string(i) + " reads neighbour " + std::to_string(nb) +
                             " of another chunk without a wait");
                }
            }
        }
        for (int32_t pc : touched) covered[(size_t)pc] = -1;
    }
    if (!schedule_completes(cs)) fail("waits form a cycle: some entries never run");
    (void)c2d;
    return schedule_stats(cs, threads);
}

bool schedule_completes(const ChunkSchedule &cs) {
    // round robin until nothing moves; all must finish
    const int32_t C = cs.n_chunks;
    const int64_t E = (int64_t)cs.ent_level.size();
    std::vector<int32_t> cur(cs.ch_ent.begin(), cs.ch_ent.end() - 1), progress((size_t)C, 0);
    int64_t done = 0;
    for (bool moved = true; moved;) {
        moved = false;
        for (int32_t w = 0; w < C; ++w) {
            while (cur[(size_t)w] < cs.ch_ent[(size_t)w + 1]) {
                const int32_t en = cur[(size_t)w];
                bool ready = true;
                for (int32_t i = cs.ent_wait[(size_t)en]; i < cs.ent_wait[(size_t)en + 1] && ready; ++i)
                    ready = progress[(size_t)cs.wait_chunk[(size_t)i]] >= cs.wait_level[(size_t)i] + 1;
                if (!ready) break;
                if (cs.ent_level[(size_t)en] & kSchedPublish) progress[(size_t)w] = (cs.ent_level[(size_t)en] & kLevelMask) + 1;
                cur[(size_t)w]++;
                done++;
                moved = true;
            }
        }
    }
    return done == E;
}

}  // namespace psst
