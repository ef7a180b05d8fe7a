// ordered_sched.cpp -- levels and chunk schedule of the reference-ordered
// time step (see ordered_sched.hpp).
#include "ordered_sched.hpp"

#include <algorithm>
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

ChunkSchedule chunk_schedule(int32_t N, int32_t chunk_cells, const std::vector<int32_t> &lev,
                             const int32_t *list_offsets, const int32_t *lnb, const std::vector<int32_t> &c2d,
                             const std::vector<int32_t> &d2c) {
    if (chunk_cells <= 0) throw std::invalid_argument("chunk_cells must be positive");
    ChunkSchedule cs;
    const int32_t K = chunk_cells;
    cs.chunk_cells = K;
    cs.n_chunks = (N + K - 1) / K;
    cs.pos_cell.resize((size_t)N);
    cs.ch_ent.assign((size_t)cs.n_chunks + 1, 0);
    cs.ent_pos.clear();
    cs.ent_level.clear();
    // positions: chunk by chunk, then by level (counting sort), then by device id
    std::vector<int32_t> count;
    int32_t pos = 0;
    for (int32_t w = 0; w < cs.n_chunks; ++w) {
        const int32_t b = w * K, e = std::min(N, b + K);
        int32_t lo = INT32_MAX, hi = -1;
        for (int32_t d = b; d < e; ++d) {
            const int32_t l = lev[(size_t)d2c[(size_t)d]];
            lo = std::min(lo, l);
            hi = std::max(hi, l);
        }
        count.assign((size_t)(hi - lo + 2), 0);
        for (int32_t d = b; d < e; ++d) count[(size_t)(lev[(size_t)d2c[(size_t)d]] - lo + 1)]++;
        for (int32_t l = 0; l <= hi - lo; ++l) {
            if (count[(size_t)l + 1] > 0) {
                cs.ent_pos.push_back(pos + count[(size_t)l]);
                cs.ent_level.push_back(lo + l);
            }
            count[(size_t)l + 1] += count[(size_t)l];
        }
        for (int32_t d = b; d < e; ++d) cs.pos_cell[(size_t)(pos + count[(size_t)(lev[(size_t)d2c[(size_t)d]] - lo)]++)] = d;
        pos += e - b;
        cs.ch_ent[(size_t)w + 1] = (int32_t)cs.ent_level.size();
    }
    cs.ent_pos.push_back(N);
    // waits: per entry, for each other chunk owning an earlier neighbour of
    // one of its cells, the largest level among those neighbours; left out
    // when an earlier entry of the chunk already waited for that level or more
    const int32_t E = (int32_t)cs.ent_level.size();
    cs.ent_wait.assign((size_t)E + 1, 0);
    std::vector<std::pair<int32_t, int32_t>> need, waited;  // (chunk, level)
    for (int32_t w = 0; w < cs.n_chunks; ++w) {
        waited.clear();
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            need.clear();
            for (int32_t p = cs.ent_pos[(size_t)en]; p < cs.ent_pos[(size_t)en + 1]; ++p) {
                const int32_t i = d2c[(size_t)cs.pos_cell[(size_t)p]];
                for (int32_t k = list_offsets[i]; k < list_offsets[i + 1]; ++k) {
                    const int32_t nb = lnb[k];
                    if (nb >= i) continue;
                    const int32_t pc = c2d[(size_t)nb] / K;
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
    // the waited-for entries publish
    for (size_t i = 0; i < cs.wait_chunk.size(); ++i) {
        const int32_t pc = cs.wait_chunk[i], L = cs.wait_level[i];
        auto b = cs.ent_level.begin() + cs.ch_ent[(size_t)pc], e = cs.ent_level.begin() + cs.ch_ent[(size_t)pc + 1];
        auto it = std::lower_bound(b, e, L, [](int32_t a, int32_t l) { return (a & (kSchedPublish - 1)) < l; });
        if (it == e || (*it & (kSchedPublish - 1)) != L) throw std::logic_error("chunk schedule: waited level missing");
        *it |= kSchedPublish;
    }
    return cs;
}

namespace {

void fail(const std::string &what) { throw std::logic_error("chunk schedule: " + what); }

}  // namespace

ScheduleStats check_schedule(const ChunkSchedule &cs, int32_t N, const std::vector<int32_t> &lev,
                             const int32_t *list_offsets, const int32_t *lnb, const std::vector<int32_t> &c2d,
                             const std::vector<int32_t> &d2c, int threads) {
    constexpr int32_t kLevelMask = kSchedPublish - 1;
    const int32_t K = cs.chunk_cells, C = cs.n_chunks;
    const int32_t E = (int32_t)cs.ent_level.size();
    if (K <= 0 || C != (N + K - 1) / K) fail("chunk count");
    if ((int32_t)cs.ch_ent.size() != C + 1 || cs.ch_ent[0] != 0 || cs.ch_ent[(size_t)C] != E) fail("ch_ent");
    if ((int32_t)cs.ent_pos.size() != E + 1 || cs.ent_pos[0] != 0 || cs.ent_pos[(size_t)E] != N) fail("ent_pos");
    if ((int32_t)cs.ent_wait.size() != E + 1 || cs.ent_wait[0] != 0 ||
        cs.ent_wait[(size_t)E] != (int32_t)cs.wait_chunk.size() || cs.wait_level.size() != cs.wait_chunk.size())
        fail("ent_wait");
    if ((int32_t)cs.pos_cell.size() != N) fail("pos_cell size");
    std::vector<char> seen((size_t)N, 0);
    std::vector<int32_t> chunk_of_entry((size_t)E);
    ScheduleStats st;
    st.chunks = C;
    st.entries = E;
    st.waits = (int64_t)cs.wait_chunk.size();
    int32_t maxlev = -1;
    for (int32_t w = 0; w < C; ++w) {
        if (cs.ch_ent[(size_t)w + 1] < cs.ch_ent[(size_t)w]) fail("ch_ent not monotone");
        // positions of the chunk are exactly its device cells
        const int32_t pb = cs.ch_ent[(size_t)w] < E ? cs.ent_pos[(size_t)cs.ch_ent[(size_t)w]] : N;
        if (pb != w * K) fail("chunk " + std::to_string(w) + " does not start at its first cell");
        int32_t prev = -1;
        for (int32_t en = cs.ch_ent[(size_t)w]; en < cs.ch_ent[(size_t)w + 1]; ++en) {
            chunk_of_entry[(size_t)en] = w;
            const int32_t l = cs.ent_level[(size_t)en] & kLevelMask;
            if (l <= prev) fail("levels of a chunk not rising");
            prev = l;
            maxlev = std::max(maxlev, l);
            if (cs.ent_level[(size_t)en] & kSchedPublish) st.publishes++;
            if (cs.ent_pos[(size_t)en + 1] <= cs.ent_pos[(size_t)en]) fail("empty entry");
            if (cs.ent_wait[(size_t)en + 1] < cs.ent_wait[(size_t)en]) fail("ent_wait not monotone");
            st.widest_entry = std::max<int64_t>(st.widest_entry, cs.ent_pos[(size_t)en + 1] - cs.ent_pos[(size_t)en]);
            for (int32_t p = cs.ent_pos[(size_t)en]; p < cs.ent_pos[(size_t)en + 1]; ++p) {
                const int32_t d = cs.pos_cell[(size_t)p];
                if (d < 0 || d >= N || seen[(size_t)d]) fail("position table");
                seen[(size_t)d] = 1;
                if (d / K != w) fail("cell outside its chunk");
                if (lev[(size_t)d2c[(size_t)d]] != l) fail("cell level differs from its entry");
            }
        }
    }
    st.levels = maxlev + 1;
    // every wait is for a level its chunk publishes at or after
    auto publish_entry = [&](int32_t pc, int32_t L) {
        for (int32_t en = cs.ch_ent[(size_t)pc]; en < cs.ch_ent[(size_t)pc + 1]; ++en)
            if ((cs.ent_level[(size_t)en] & kSchedPublish) && (cs.ent_level[(size_t)en] & kLevelMask) >= L) return en;
        return -1;
    };
    std::vector<int32_t> wait_src(cs.wait_chunk.size());
    for (size_t i = 0; i < cs.wait_chunk.size(); ++i) {
        const int32_t pc = cs.wait_chunk[i];
        if (pc < 0 || pc >= C) fail("waited chunk out of range");
        wait_src[i] = publish_entry(pc, cs.wait_level[i]);
        if (wait_src[i] < 0) fail("wait for a level that is never published");
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
                    const int32_t pc = c2d[(size_t)nb] / K;
                    if (pc != w && covered[(size_t)pc] < lev[(size_t)nb])
                        fail("cell " + std::to_string(i) + " reads neighbour " + std::to_string(nb) +
                             " of another chunk without a wait");
                }
            }
        }
        for (int32_t pc : touched) covered[(size_t)pc] = -1;
    }
    // run as workgroups: round robin until nothing moves; all must finish
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
    if (done != E) fail("waits form a cycle: " + std::to_string(E - done) + " entries never run");
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
                start = std::max(start, fin[(size_t)wait_src[(size_t)i]] + wait_cost);
            const int64_t cells = cs.ent_pos[(size_t)en + 1] - cs.ent_pos[(size_t)en];
            fin[(size_t)en] = start + (cells + threads - 1) / threads;
            span = std::max(span, fin[(size_t)en]);
        }
        (wait_cost == 0 ? st.makespan_free : st.makespan_wait4) = span;
    }
    return st;
}

}  // namespace psst
