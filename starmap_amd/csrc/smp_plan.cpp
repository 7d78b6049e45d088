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

std::string plan_adjoint(int64_t n_steps, int every, int64_t buffers, AdjSchedule &out) {
    out = AdjSchedule();
    if (n_steps < 1) return "n_steps must be >= 1";
    if (every < 1) return "checkpoint_every must be >= 1";
    const int64_t k = std::min<int64_t>(every, n_steps);
    const int64_t U64 = (n_steps + k - 1) / k;
    if (U64 > 1 << 20) return "too many segments: raise checkpoint_every";
    const int U = (int)U64;
    const int64_t rb = k + 1;   // buffers of one record range
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
