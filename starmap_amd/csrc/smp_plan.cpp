// Host-side planning for the StaRMAP substep kernels.  See smp_plan.h.
#include "smp_plan.h"

#include <algorithm>
#include <cstring>
#include <deque>
#include <limits>
#include <set>
#include <sstream>

namespace smp {

std::string build_geometry(Geometry &g) {
    if (g.n_sys <= 0) return "n_sys must be positive";
    for (int d = 0; d < 3; ++d) {
        if (g.n[d] < 1) return "every n(d) must be >= 1";
        if (g.bc[d] != 0 && g.bc[d] != 1) return "bc(d) must be 0 (periodic) or 1 (extrapolation)";
        if (!(g.h[d] > 0)) return "every h(d) must be positive";
        const Csr &M = g.M[d];
        if (M.rp.empty()) continue;
        if ((int)M.rp.size() != g.n_sys + 1) return "row_ptr must have n_sys+1 entries";
        if (M.rp[0] != 0 || M.rp[g.n_sys] != (int)M.ci.size()) return "row_ptr inconsistent with col";
        for (int j = 0; j < g.n_sys; ++j) {
            if (M.rp[j + 1] < M.rp[j]) return "row_ptr not non-decreasing";
            for (int p = M.rp[j]; p < M.rp[j + 1]; ++p) {
                if (M.ci[p] < 0 || M.ci[p] >= g.n_sys) return "column index out of range";
                if (p > M.rp[j] && M.ci[p] <= M.ci[p - 1]) return "columns must ascend within a row";
            }
        }
    }
    // Grid types: component 1 lives on the (1,1,1) grid; a coupling along
    // axis d connects components whose grid types differ exactly in d.  The
    // reference reaches the same assignment by its fixed-point iteration
    // over the eight classes (starmap_solver.m:80-98).
    g.gt.assign(3 * g.n_sys, 0);
    std::vector<char> seen(g.n_sys, 0);
    std::deque<int> queue;
    g.gt[0] = g.gt[1] = g.gt[2] = 1;
    seen[0] = 1;
    queue.push_back(0);
    while (!queue.empty()) {
        int j = queue.front();
        queue.pop_front();
        for (int d = 0; d < 3; ++d) {
            const Csr &M = g.M[d];
            if (M.rp.empty()) continue;
            for (int p = M.rp[j]; p < M.rp[j + 1]; ++p) {
                int k = M.ci[p];
                int want[3] = {g.gt[3 * j], g.gt[3 * j + 1], g.gt[3 * j + 2]};
                want[d] = 3 - want[d];
                if (seen[k]) {
                    for (int e = 0; e < 3; ++e)
                        if (g.gt[3 * k + e] != want[e]) {
                            std::ostringstream os;
                            os << "coupling is not bipartite: components " << j << " and " << k
                               << " along axis " << d;
                            return os.str();
                        }
                } else {
                    for (int e = 0; e < 3; ++e) g.gt[3 * k + e] = want[e];
                    seen[k] = 1;
                    queue.push_back(k);
                }
            }
        }
    }
    for (int j = 0; j < g.n_sys; ++j)
        if (!seen[j]) {
            std::ostringstream os;
            os << "component " << j << " is not coupled to component 0";
            return os.str();
        }
    // kernel axes: solver axes that carry a coupling or more than one cell
    g.ndim = 0;
    for (int d = 0; d < 3; ++d)
        if (!g.M[d].empty() || g.n[d] > 1) g.axis[g.ndim++] = d;
    if (g.ndim == 0) g.axis[g.ndim++] = 2;
    for (int e = g.ndim; e < 3; ++e) g.axis[e] = -1;
    g.even.clear();
    g.odd.clear();
    for (int j = 0; j < g.n_sys; ++j) {
        int par = (g.gt[3 * j] - 1) + (g.gt[3 * j + 1] - 1) + (g.gt[3 * j + 2] - 1);
        (par % 2 ? g.odd : g.even).push_back(j);
    }
    return "";
}

std::string make_tile(const Geometry &g, int tx, int ty, int tz, TileShape &ts) {
    ts = TileShape();
    const int t[3] = {tx, ty, tz};
    for (int e = 0; e < 3; ++e) {
        ts.tile[e] = e < g.ndim ? t[e] : 1;
        if (ts.tile[e] < 1) return "tile extents must be positive";
    }
    // LDS rows of a power of two: the tile plus its one-cell halo
    if ((ts.tile[0] + 1) & ts.tile[0]) return "tile[0]+1 must be a power of two";
    ts.threads = 256;
    if (ts.cells() > 256) {
        if (ts.cells() > 512 || ts.tile[0] < 3 || ((ts.tile[0] + 1) / 2) * ts.tile[1] * ts.tile[2] > 256)
            return "a tile holds at most 256 cells, or 512 with (tile[0]+1)/2*tile[1]*tile[2] <= 256";
        ts.pair = true;
        ts.threads = 512;
    }
    for (int e = 0; e < 3; ++e) ts.lds[e] = e < g.ndim ? ts.tile[e] + 1 : 1;
    ts.slot_elems = ts.lds[0] * ts.lds[1] * ts.lds[2];
    return "";
}

TileShape choose_tile(const Geometry &g, int elem_bytes) {
    (void)elem_bytes;
    // the first axis spans a power of two with its halo (LDS rows of 256, 64
    // or 32 elements); 256 threads, one per cell, spare threads only copy
    TileShape ts;
    if (g.ndim <= 1)
        make_tile(g, 255, 1, 1, ts);
    else if (g.ndim == 2)
        make_tile(g, 63, 4, 1, ts);
    else
        make_tile(g, 31, 4, 2, ts);
    return ts;
}

std::string build_plan(const Geometry &g, const TileShape &ts, int kind, int slot_capacity,
                       int group_size, int max_load_elems, int max_loads_per_group, Plan &out) {
    out = Plan();
    const std::vector<int32_t> &tg = kind == 0 ? g.odd : g.even;
    int kaxis_of[3] = {-1, -1, -1};
    for (int e = 0; e < g.ndim; ++e) kaxis_of[g.axis[e]] = e;
    const int stride[3] = {1, ts.lds[0], ts.lds[0] * ts.lds[1]};
    const int max_loads = std::max(1, std::min(max_loads_per_group, max_load_elems / ts.slot_elems));

    // sources of every target
    std::vector<std::vector<int>> src(tg.size());
    std::set<int> all_src;
    for (size_t t = 0; t < tg.size(); ++t) {
        std::set<int> s;
        int nt = 0;
        for (int d = 0; d < 3; ++d) {
            const Csr &M = g.M[d];
            if (M.rp.empty()) continue;
            nt += M.rp[tg[t] + 1] - M.rp[tg[t]];
            for (int p = M.rp[tg[t]]; p < M.rp[tg[t] + 1]; ++p) s.insert(M.ci[p]);
        }
        if (nt > MAXT) return "a coupling row has more than 12 nonzeros";
        out.max_terms = std::max(out.max_terms, nt);
        src[t].assign(s.begin(), s.end());
        all_src.insert(s.begin(), s.end());
        if ((int)s.size() > slot_capacity) return "a target couples to more components than fit in LDS";
    }
    out.n_sources = (int)all_src.size();

    // groups of consecutive targets; a group's tiles are copied in while the
    // previous group is still read, so two consecutive groups must fit
    // together: each group gets half of the slots
    const int group_cap = (slot_capacity - 1) / 2;
    std::vector<std::pair<int, int>> groups;
    std::vector<std::vector<int>> gsrc;
    for (size_t t = 0; t < tg.size();) {
        size_t t1 = t;
        std::set<int> s;
        while (t1 < tg.size() && (int)(t1 - t) < group_size) {
            std::set<int> s2 = s;
            s2.insert(src[t1].begin(), src[t1].end());
            if ((int)s2.size() > group_cap) break;
            s = s2;
            ++t1;
        }
        if (t1 == t) return "a target couples to more components than half the LDS slots";
        groups.push_back({(int)t, (int)t1});
        gsrc.push_back(std::vector<int>(s.begin(), s.end()));
        t = t1;
    }

    // slot assignment: keep what the next groups need, evict the source
    // whose next use is farthest away
    const int INF = std::numeric_limits<int>::max();
    std::vector<int> slot_of(g.n_sys, -1);
    // slot 0 is a zero tile: padding terms read it
    std::vector<int> owner(slot_capacity, -1);
    if (slot_capacity < 2) return "too few LDS slots";
    owner[0] = -2;
    auto next_use = [&](int k, size_t from) {
        for (size_t q = from; q < gsrc.size(); ++q)
            if (std::binary_search(gsrc[q].begin(), gsrc[q].end(), k)) return (int)q;
        return INF;
    };
    int used = 0;
    for (size_t q = 0; q < groups.size(); ++q) {
        std::vector<DevLoad> new_loads;
        for (int k : gsrc[q]) {
            if (slot_of[k] >= 0) continue;
            int s = -1;
            for (int c = 1; c < slot_capacity; ++c)
                if (owner[c] == -1) { s = c; break; }
            if (s < 0) {
                int best = -1, far = -1;
                for (int c = 1; c < slot_capacity; ++c) {
                    int o = owner[c];
                    if (std::binary_search(gsrc[q].begin(), gsrc[q].end(), o)) continue;
                    // the previous group still reads its sources while this
                    // group's tiles are copied in
                    if (q > 0 && std::binary_search(gsrc[q - 1].begin(), gsrc[q - 1].end(), o))
                        continue;
                    int nu = next_use(o, q + 1);
                    if (nu > far) { far = nu; best = c; }
                }
                if (best < 0) return "LDS slot assignment failed";
                slot_of[owner[best]] = -1;
                s = best;
            }
            owner[s] = k;
            slot_of[k] = s;
            used = std::max(used, s + 1);
            DevLoad ld;
            std::memset(&ld, 0, sizeof(ld));
            ld.comp = k;
            ld.slot = s;
            for (int e = 0; e < 3; ++e)
                ld.org[e] = e < g.ndim && g.gt[3 * k + g.axis[e]] == 1 ? -1 : 0;
            new_loads.push_back(ld);
        }
        // staging is bounded by the prefetch registers: leading load-only
        // groups take what does not fit
        size_t first = 0;
        while (new_loads.size() - first > (size_t)max_loads) {
            DevGroup lg;
            lg.l0 = (int)out.loads.size();
            out.loads.insert(out.loads.end(), new_loads.begin() + first,
                             new_loads.begin() + first + max_loads);
            lg.l1 = (int)out.loads.size();
            lg.t0 = lg.t1 = (int)out.targets.size();
            out.groups.push_back(lg);
            first += max_loads;
        }
        DevGroup gr;
        gr.l0 = (int)out.loads.size();
        out.loads.insert(out.loads.end(), new_loads.begin() + first, new_loads.end());
        gr.l1 = (int)out.loads.size();
        gr.t0 = (int)out.targets.size();
        for (int t = groups[q].first; t < groups[q].second; ++t) {
            int j = tg[t];
            TgtRec r;
            std::memset(&r, 0, sizeof(r));
            r.comp = j;
            for (int e = 0; e < 3; ++e) r.ext[e] = e < g.ndim ? g.ext_of(j, g.axis[e]) : 1;
            // x-terms, then y-terms, then z-terms, ascending k within each:
            // the summation order of sumcell (starmap_solver.m:350-351, :516-519)
            int nt = 0;
            for (int d = 0; d < 3; ++d) {
                const Csr &M = g.M[d];
                if (M.rp.empty()) continue;
                int kd = kaxis_of[d];
                for (int p = M.rp[j]; p < M.rp[j + 1]; ++p) {
                    int k = M.ci[p];
                    if (kd < 0) return "coupling along an axis without cells";
                    int off = slot_of[k] * ts.slot_elems;
                    for (int e = 0; e < g.ndim; ++e)
                        if (e != kd && g.gt[3 * k + g.axis[e]] == 1) off += stride[e];
                    r.t[nt].off = off;
                    r.t[nt].stride = stride[kd];
                    r.t[nt].coef = -M.cv[p] / g.h[d];
                    ++nt;
                }
            }
            // padding terms (coefficient 0 times a difference within the zero
            // tile of slot 0) keep the kernel's term loop free of branches
            for (int q = nt; q < MAXT; ++q) {
                r.t[q].off = 0;
                r.t[q].stride = 0;
                r.t[q].coef = 0.0;
            }
            out.targets.push_back(r);
        }
        gr.t1 = (int)out.targets.size();
        out.groups.push_back(gr);
    }
    out.n_slots = used;
    return "";
}

}  // namespace smp
