// Host-side planning for the StaRMAP substep kernels.  See smp_plan.h.
#include "smp_plan.h"

#include <algorithm>
#include <climits>
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

namespace {

// One group of a slot schedule: targets [t0, t1), the source tiles it
// stages (source id, slot), and the slot of each of its sources (ids sorted)
struct SlotGroup {
    int t0 = 0, t1 = 0;
    std::vector<std::pair<int, int>> loads;
    std::vector<int> ids, slots;
    int slot(int id) const {
        const auto it = std::lower_bound(ids.begin(), ids.end(), id);
        return slots[it - ids.begin()];
    }
};

// Groups of consecutive targets and their LDS slots.  A group's tiles are
// copied in while the previous group is still read, so two consecutive
// groups must fit together: each group gets half of the slots.  Slot 0 is a
// zero tile (padding terms read it).  A new tile takes a free slot, or
// evicts the source whose next use is farthest away, never one the current
// or the previous group reads.
// fixed_groups: every group takes exactly group_size consecutive targets
// (an error if their sources do not fit)
std::string schedule_slots(const std::vector<std::vector<int>> &src, int n_ids, int slot_capacity,
                           int group_size, std::vector<SlotGroup> &out, int &used,
                           bool fixed_groups = false) {
    out.clear();
    used = 0;
    if (slot_capacity < 2) return "too few LDS slots";
    const int group_cap = (slot_capacity - 1) / 2;
    std::vector<std::vector<int>> gsrc;
    for (size_t t = 0; t < src.size();) {
        size_t t1 = t;
        std::set<int> s;
        while (t1 < src.size() && (int)(t1 - t) < group_size) {
            std::set<int> s2 = s;
            s2.insert(src[t1].begin(), src[t1].end());
            if ((int)s2.size() > group_cap) break;
            s = s2;
            ++t1;
        }
        if (t1 == t) return "a target couples to more components than half the LDS slots";
        if (fixed_groups && (int)(t1 - t) < group_size && t1 < src.size())
            return "a group's sources do not fit in half the LDS slots";
        SlotGroup gr;
        gr.t0 = (int)t;
        gr.t1 = (int)t1;
        out.push_back(gr);
        gsrc.push_back(std::vector<int>(s.begin(), s.end()));
        t = t1;
    }
    const int INF = std::numeric_limits<int>::max();
    std::vector<int> slot_of(n_ids, -1);
    std::vector<int> owner(slot_capacity, -1);
    owner[0] = -2;
    auto next_use = [&](int k, size_t from) {
        for (size_t q = from; q < gsrc.size(); ++q)
            if (std::binary_search(gsrc[q].begin(), gsrc[q].end(), k)) return (int)q;
        return INF;
    };
    for (size_t q = 0; q < out.size(); ++q) {
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
            out[q].loads.push_back({k, s});
        }
        out[q].ids = gsrc[q];
        for (int k : gsrc[q]) out[q].slots.push_back(slot_of[k]);
    }
    return "";
}

// LDS element offset of the low sample of a term of a target from source
// tile `slot` of a component of grid type gts along kernel axis kd: the
// slot base, shifted by one on every other axis where the source is
// cell-centred (its slot starts one cell lower)
int term_offset(const Geometry &g, const TileShape &ts, int slot, const int32_t *gts, int kd) {
    const int stride[3] = {1, ts.lds[0], ts.lds[0] * ts.lds[1]};
    int off = slot * ts.slot_elems;
    for (int e = 0; e < g.ndim; ++e)
        if (e != kd && gts[g.axis[e]] == 1) off += stride[e];
    return off;
}

DevLoad load_record(const Geometry &g, int comp, int slot) {
    DevLoad ld;
    std::memset(&ld, 0, sizeof(ld));
    ld.comp = comp;
    ld.slot = slot;
    for (int e = 0; e < 3; ++e) ld.org[e] = e < g.ndim && g.gt[3 * comp + g.axis[e]] == 1 ? -1 : 0;
    return ld;
}

}  // namespace

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

    std::vector<SlotGroup> sg;
    int used = 0;
    std::string err = schedule_slots(src, g.n_sys, slot_capacity, group_size, sg, used);
    if (!err.empty()) return err;
    for (const SlotGroup &q : sg) {
        std::vector<DevLoad> new_loads;
        for (const auto &ks : q.loads) new_loads.push_back(load_record(g, ks.first, ks.second));
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
        for (int t = q.t0; t < q.t1; ++t) {
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
                    r.t[nt].off = term_offset(g, ts, q.slot(k), &g.gt[3 * k], kd);
                    r.t[nt].stride = stride[kd];
                    r.t[nt].coef = -M.cv[p] / g.h[d];
                    ++nt;
                }
            }
            // padding terms (coefficient 0 times a difference within the zero
            // tile of slot 0) keep the kernel's term loop free of branches
            for (int p = nt; p < MAXT; ++p) {
                r.t[p].off = 0;
                r.t[p].stride = 0;
                r.t[p].coef = 0.0;
            }
            out.targets.push_back(r);
        }
        gr.t1 = (int)out.targets.size();
        out.groups.push_back(gr);
    }
    out.n_slots = used;
    return "";
}

std::string build_adj_plan(const Geometry &g, const TileShape &ts, int kind, int slot_capacity,
                           bool with_w, AdjPlan &out) {
    out = AdjPlan();
    if (ts.pair) return "the adjoint update takes one cell per thread";
    if (ADJ_TERMS * g.ndim > MAXT) return "too many kernel axes";
    const int n = g.n_sys;
    int kaxis_of[3] = {-1, -1, -1};
    for (int e = 0; e < g.ndim; ++e) kaxis_of[g.axis[e]] = e;
    // targets of the parity grouped by grid type (codes in order of first
    // appearance), components ascending within a grid type
    const std::vector<int32_t> &cs = kind == 0 ? g.odd : g.even;
    auto code_of = [&](int j) { return (g.gt[3 * j] - 1) * 4 + (g.gt[3 * j + 1] - 1) * 2 + (g.gt[3 * j + 2] - 1); };
    for (int j : cs)
        if (std::find(out.gcode.begin(), out.gcode.end(), code_of(j)) == out.gcode.end())
            out.gcode.push_back(code_of(j));
    std::vector<int> tg;
    for (int c : out.gcode)
        for (int j : cs)
            if (code_of(j) == c) tg.push_back(j);
    // transposed coupling lists: rows i with M_d(i, j) != 0, ascending i
    std::vector<std::vector<std::vector<std::pair<int, double>>>> cols(
        3, std::vector<std::vector<std::pair<int, double>>>(n));
    for (int d = 0; d < 3; ++d) {
        const Csr &M = g.M[d];
        if (M.rp.empty()) continue;
        for (int i = 0; i < n; ++i)
            for (int p = M.rp[i]; p < M.rp[i + 1]; ++p) cols[d][M.ci[p]].push_back({i, M.cv[p]});
    }
    // source ids: k the state tile of component k; n*(1+m) + i the mu tile
    // of component i, m = 0 plain, m = 1+e with the boundary faces of kernel
    // axis e zero (a coupling along e to a target cell-centred on e)
    auto mu_id = [&](int j, int i, int e) {
        const bool masked = g.bc[g.axis[e]] == 1 && g.gt[3 * j + g.axis[e]] == 1;
        return n * (1 + (masked ? 1 + e : 0)) + i;
    };
    std::vector<std::vector<int>> src(tg.size());
    for (size_t t = 0; t < tg.size(); ++t) {
        const int j = tg[t];
        std::set<int> s;
        for (int d = 0; d < 3; ++d) {
            const Csr &M = g.M[d];
            if (M.rp.empty()) continue;
            const int e = kaxis_of[d];
            if (e < 0) return "coupling along an axis without cells";
            if (M.rp[j + 1] - M.rp[j] > ADJ_TERMS || (int)cols[d][j].size() > ADJ_TERMS)
                return "more than 4 couplings of a component along an axis";
            if (!with_w)
                for (int p = M.rp[j]; p < M.rp[j + 1]; ++p) s.insert(M.ci[p]);
            for (const auto &ic : cols[d][j]) s.insert(mu_id(j, ic.first, e));
        }
        src[t].assign(s.begin(), s.end());
    }
    std::vector<SlotGroup> sg;
    int used = 0;
    std::string err = schedule_slots(src, 5 * n, slot_capacity, ADJ_RECS, sg, used);
    if (!err.empty()) return err;
    // the record positions of each group: the group's targets in order
    // (positions p*ADJ_TPH .. (p+1)*ADJ_TPH-1 for part p of the cell
    // threads), then positions without a target
    std::vector<int> pos_of;   // record index of every target
    for (const SlotGroup &q : sg) {
        if ((int)q.loads.size() > ADJ_MAXL) return "more than 64 source tiles in a group";
        DevGroup gr;
        gr.l0 = (int)out.loads.size();
        for (const auto &ks : q.loads) {
            const int mode = ks.first / n;   // 0 state, 1 mu, 2+e mu masked on e
            DevLoad ld = load_record(g, ks.first % n, ks.second);
            ld.pad[0] = mode;
            out.loads.push_back(ld);
        }
        gr.l1 = (int)out.loads.size();
        gr.t0 = (int)out.targets.size();
        for (int p = 0; p < ADJ_RECS; ++p) {
            AdjRec r;
            std::memset(&r, 0, sizeof(r));
            r.comp = -1;
            const int t = q.t0 + p;
            if (t < q.t1) {
                const int j = tg[t];
                r.comp = j;
                r.gq = (int32_t)(std::find(out.gcode.begin(), out.gcode.end(), code_of(j)) - out.gcode.begin());
                r.jpos = j > 0;
                for (int e = 0; e < 3; ++e) r.ext[e] = e < g.ndim ? g.ext_of(j, g.axis[e]) : 1;
                for (int e = 0; e < g.ndim; ++e) {
                    const int d = g.axis[e];
                    const Csr &M = g.M[d];
                    if (M.rp.empty()) continue;
                    int nf = 0, nb = 0;
                    for (int pp = M.rp[j]; pp < M.rp[j + 1] && !with_w; ++pp, ++nf) {
                        const int k = M.ci[pp];
                        r.off[ADJ_TERMS * e + nf] = term_offset(g, ts, q.slot(k), &g.gt[3 * k], e);
                        r.coef[ADJ_TERMS * e + nf] = -M.cv[pp] / g.h[d];
                    }
                    // lambda_j += sum_i c_ij (D^T mu_i): the forward-form
                    // difference of mu_i at j's points, coefficient -c_ij
                    for (const auto &ic : cols[d][j]) {
                        const int i = ic.first, qq = MAXT + ADJ_TERMS * e + nb++;
                        r.off[qq] = term_offset(g, ts, q.slot(mu_id(j, i, e)), &g.gt[3 * i], e);
                        r.coef[qq] = ic.second / g.h[d];
                    }
                }
                pos_of.push_back((int)out.targets.size());
            }
            out.targets.push_back(r);
        }
        gr.t1 = (int)out.targets.size();
        out.groups.push_back(gr);
    }
    // each part sums the gradient of a grid type over its targets of that
    // grid type in order and stores the sum after the last of them
    for (int h = 0; h < ADJ_PARTS; ++h) {
        int last = -1;
        for (int p : pos_of) {
            if ((p % ADJ_RECS) / ADJ_TPH != h) continue;
            if (last >= 0 && out.targets[last].gq != out.targets[p].gq) out.targets[last].flush = 1;
            last = p;
        }
        if (last >= 0) out.targets[last].flush = 1;
    }
    out.n_slots = used;
    return "";
}

// ---------------------------------------------------------------------------
// Adjoint schedule (see smp_plan.h)

namespace {

// Extra segment advances of reversing l segments from a checkpoint at their
// start with f further checkpoint slots: N (the segments were not passed by
// the forward pass in order) and S (they lie on the forward pass's way, so
// advancing through them from the left is free).  A split at m stores the
// state after m segments in a slot, reverses the right part with one slot
// less, frees the slot and reverses the left part.  Without a free slot
// every segment is re-advanced from the checkpoint.
struct AdjCost {
    int L = 0, F = 0;
    std::vector<int64_t> N, S;
    std::vector<int32_t> mN, mS;
    int64_t &n(int l, int f) { return N[(size_t)l * (F + 1) + f]; }
    int64_t &s(int l, int f) { return S[(size_t)l * (F + 1) + f]; }
    int32_t &sn(int l, int f) { return mN[(size_t)l * (F + 1) + f]; }
    int32_t &ss(int l, int f) { return mS[(size_t)l * (F + 1) + f]; }
    void build(int l_max, int f_max) {
        L = l_max;
        F = f_max;
        const size_t sz = (size_t)(L + 1) * (F + 1);
        N.assign(sz, 0);
        S.assign(sz, 0);
        mN.assign(sz, 0);
        mS.assign(sz, 0);
        for (int l = 1; l <= L; ++l)
            for (int f = 0; f <= F; ++f) {
                if (l == 1) continue;
                if (f == 0) {
                    n(l, f) = s(l, f) = (int64_t)l * (l - 1) / 2;
                    continue;
                }
                if (f >= l - 1) {
                    // a slot at every segment start on the way
                    n(l, f) = l - 1;
                    s(l, f) = 0;
                    sn(l, f) = ss(l, f) = 1;
                    continue;
                }
                int64_t bn = INT64_MAX, bs = INT64_MAX;
                for (int m = 1; m < l; ++m) {
                    const int64_t cn = m + n(l - m, f - 1) + n(m, f);
                    const int64_t cs = s(l - m, f - 1) + n(m, f);
                    if (cn < bn) { bn = cn; sn(l, f) = m; }
                    if (cs < bs) { bs = cs; ss(l, f) = m; }
                }
                n(l, f) = bn;
                s(l, f) = bs;
            }
    }
};

struct AdjEmitter {
    AdjSchedule &out;
    AdjCost &cost;
    std::vector<int> free_slots;
    int head_reversed = 0;
    int hr[2] = {0, 1};
    std::vector<AdjAction> head_rev;

    AdjEmitter(AdjSchedule &o, AdjCost &c) : out(o), cost(c) {}
    static AdjAction act(int kind, int s0, int s1, int src, int dst, int range) {
        AdjAction a;
        a.kind = kind;
        a.s0 = s0;
        a.s1 = s1;
        a.src = src;
        a.dst = dst;
        a.range = range;
        return a;
    }
    int take_slot() {
        const int s = free_slots.back();
        free_slots.pop_back();
        return s;
    }
    void reverse_one(int seg, int src) {
        const int r = hr[head_reversed++ % 2];
        head_rev.push_back(act(ADJ_RECOMP, seg, seg + 1, src, ADJ_NONE, r));
        head_rev.push_back(act(ADJ_REVERSE, seg, seg + 1, src, ADJ_NONE, r));
    }
    // reverse segments [a, b) from the checkpoint of a in slot sa, with f
    // free slots; spine: the forward pass still has to pass through them
    void rh(int a, int b, int sa, int f, bool spine) {
        const int l = b - a;
        if (l == 1) {
            if (spine) out.fwd.push_back(act(ADJ_FWD_RUN, a, b, ADJ_NONE, ADJ_NONE, ADJ_NONE));
            reverse_one(a, sa);
            return;
        }
        if (f == 0) {
            if (spine) out.fwd.push_back(act(ADJ_FWD_RUN, a, b, ADJ_NONE, ADJ_NONE, ADJ_NONE));
            for (int sg = b - 1; sg >= a; --sg) {
                int src = sa;
                if (sg > a) {
                    head_rev.push_back(act(ADJ_ADVANCE, a, sg, sa, ADJ_WORK, ADJ_NONE));
                    out.extra_advances += sg - a;
                    out.work = 1;
                    src = ADJ_WORK;
                }
                reverse_one(sg, src);
            }
            return;
        }
        // with a slot for every segment start, store each one on the way
        const int m = f >= l - 1 ? a + 1 : a + (spine ? cost.ss(l, f) : cost.sn(l, f));
        const int sm = take_slot();
        if (spine) {
            out.fwd.push_back(act(ADJ_FWD_RUN, a, m, ADJ_NONE, ADJ_NONE, ADJ_NONE));
            out.fwd.push_back(act(ADJ_FWD_CKPT, m, m, ADJ_NONE, sm, ADJ_NONE));
        } else {
            head_rev.push_back(act(ADJ_ADVANCE, a, m, sa, sm, ADJ_NONE));
            out.extra_advances += m - a;
        }
        rh(m, b, sm, f - 1, spine);
        free_slots.push_back(sm);
        rh(a, m, sa, f, false);
    }
};

}  // namespace

std::string plan_adjoint(int64_t n_steps, int every, int64_t buffers, int rec_factor,
                         AdjSchedule &out) {
    out = AdjSchedule();
    if (n_steps < 1) return "n_steps must be >= 1";
    if (every < 1) return "checkpoint_every must be >= 1";
    if (rec_factor < 1 || rec_factor > 2) return "rec_factor must be 1 or 2";
    const int64_t k = std::min<int64_t>(every, n_steps);
    const int64_t U64 = (n_steps + k - 1) / k;
    if (U64 > 1 << 20) return "too many segments: raise checkpoint_every";
    const int U = (int)U64;
    const int64_t rb = rec_factor * (k + 1);   // buffers of one record range
    out.rec_factor = rec_factor;
    out.segments = U;
    out.seg_steps = (int32_t)k;
    auto ranges_of = [&](int T, int H) -> int {
        if (H == 0) return T;
        return T >= 1 ? std::max(T, 3) : 2;
    };
    int best_T = -1, best_c = 0;
    int64_t best_cost = INT64_MAX;
    AdjCost cost;
    if (buffers < 0) {
        best_T = 0;
        best_c = U;
        best_cost = U;
    } else {
        // the dynamic programme covers head ranges of up to 4096 segments
        for (int T = U; T >= 0; --T) {
            const int H = U - T;
            if (H > 4096) break;
            const int64_t fixed = ranges_of(T, H) * rb + (T > 0 ? 1 : 0);
            if (H == 0) {
                if (fixed <= buffers) {
                    best_T = T;
                    best_c = 0;
                    best_cost = 0;
                    break;
                }
                continue;
            }
            int64_t c = std::min<int64_t>(H, buffers - fixed);
            if (c < H) c = std::min<int64_t>(H, buffers - fixed - 1);   // and the scratch state
            if (c < 1) continue;
            int64_t extra = 0;
            if (c < H) {
                if (cost.L < H || cost.F < c - 1) cost.build(H, (int)std::min<int64_t>(c - 1, H - 1));
                extra = cost.s(H, (int)std::min<int64_t>(c - 1, cost.F));
            }
            const int64_t tot = H + extra;
            if (tot < best_cost) {
                best_cost = tot;
                best_T = T;
                best_c = (int)c;
            }
        }
        if (best_T < 0) {
            std::ostringstream os;
            os << "adjoint memory budget too small: " << buffers << " state buffers, at least "
               << (std::min<int64_t>(U, 1) + 2 * rb + 2) << " needed for segments of " << k << " steps";
            return os.str();
        }
    }
    const int T = best_T, H = U - T, c = best_c;
    out.tail = T;
    out.slots = c;
    out.recomputed = H;
    out.ranges = ranges_of(T, H);
    if (H > 0 && c < H && (cost.L < H || cost.F < c - 1))
        cost.build(H, std::min(c - 1, H - 1));
    AdjEmitter em(out, cost);
    if (T >= 1) {
        em.hr[0] = 1;
        em.hr[1] = 2;
    }
    if (H > 0) {
        for (int s = c - 1; s >= 0; --s) em.free_slots.push_back(s);
        const int s0 = em.take_slot();
        out.fwd.push_back(AdjEmitter::act(ADJ_FWD_CKPT, 0, 0, ADJ_NONE, s0, ADJ_NONE));
        em.rh(0, H, s0, c - 1, true);
    }
    if (T > 0) {
        out.fwd.push_back(AdjEmitter::act(ADJ_FWD_CKPT, H, H, ADJ_NONE, ADJ_TAIL_START, ADJ_NONE));
        for (int s = H; s < U; ++s)
            out.fwd.push_back(AdjEmitter::act(ADJ_FWD_REC, s, s + 1, s == H ? ADJ_TAIL_START : ADJ_PREV,
                                              ADJ_NONE, s - H));
        out.fwd.push_back(AdjEmitter::act(ADJ_FWD_FINAL, U - 1, U, ADJ_NONE, ADJ_NONE, T - 1));
        for (int s = U - 1; s >= H; --s)
            out.rev.push_back(AdjEmitter::act(ADJ_REVERSE, s, s + 1, s == H ? ADJ_TAIL_START : ADJ_PREV,
                                              ADJ_NONE, s - H));
    }
    out.rev.insert(out.rev.end(), em.head_rev.begin(), em.head_rev.end());
    out.buffers = out.ranges * rb + (T > 0 ? 1 : 0) + c + out.work;
    return "";
}

}  // namespace smp
