// Host-side planning for the StaRMAP substep kernels (plain C++, no HIP).
//
// A substep updates every component of one grid parity ("targets") from
// difference quotients of the coupled components of the other parity
// ("sources"), starmap_solver.m:313-359 (odd targets) and :361-414 (even
// targets).  The device kernel processes one spatial tile per workgroup and
// walks a schedule of groups: while it updates the targets of group g from
// LDS, it already fetches the source tiles and target values of group g+1.
// This file builds that schedule once per problem.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace smp {

constexpr int MAXT = 12;      // coupling terms per target record

// device-visible records (POD, shared with the kernels)
struct DevGroup {
    int32_t l0, l1;     // loads [l0,l1)
    int32_t t0, t1;     // targets [t0,t1)
};

struct DevLoad {
    int32_t comp;       // source component
    int32_t slot;       // LDS slot
    int32_t org[3];     // tile origin shift per kernel axis (0 or -1)
    int32_t pad[3];
};

// One coupling term: W += coef * (lds[off + lbase + stride] - lds[off + lbase]).
// The schedule holds element offsets relative to the first source slot; the
// device copy holds byte offsets from the start of the kernel's LDS.
struct alignas(16) Term {
    int32_t off;        // offset of the low sample (slot base + shifts)
    int32_t stride;     // stride of the coupled axis
    union {
        double coef;        // -M(j,k)/h; 0 for padding terms
        float coef32[2];    // the same rounded to FP32 (FP32-state device copy)
    };
};

// One target of a substep: its coupling terms and its coefficients.  The
// coefficient part is filled in by the runtime from the current material,
// source and filter data.  A coefficient is a scalar (*_off == -1), a field
// staged per tile in LDS (*_off >= 0: offset of its tile) or a field read
// from global memory (*_off == -2, pointer *_p).
struct alignas(16) TgtRec {
    int32_t comp;
    int32_t filt;       // apply the substep filter after each application
    int32_t sig_off, E_off;
    int32_t Q_off;
    int32_t ext[3];     // target extent per kernel axis
    union {
        struct {
            const void *sig_p, *E_p, *Q_p;   // fields read from global memory (off -2)
        };
        // a record with no field read from global memory, FP32-state device
        // copy: dt/2*sigma*E and dt/2*E of scalar coefficients, 1-F of the
        // filter slots k=1,2, sigma (the FP32 half-step, smp_kernels.hip)
        float f32[6];
    };
    int64_t uoff;       // byte offset of the component in the state (device copy)
    const void *F[2];   // filter fields of slots k=1,2 (or null: F_s)
    double sig_s, E_s;
    double F_s[2];      // filter factors of slots k=1,2 (1 without a filter)
    Term t[MAXT];
};

struct Csr {
    std::vector<int32_t> rp, ci;
    std::vector<double> cv;
    bool empty() const { return ci.empty(); }
};

struct Geometry {
    int n_sys = 0;
    int n[3] = {1, 1, 1};       // solver axes
    int bc[3] = {0, 0, 0};
    double h[3] = {1, 1, 1};
    Csr M[3];                    // Mx, My, Mz
    std::vector<int32_t> gt;     // 3*n_sys, values 1/2 (solver axes)
    int ndim = 0;                // number of kernel axes
    int axis[3] = {0, 1, 2};     // kernel axis -> solver axis
    std::vector<int32_t> even, odd;
    int ext_of(int j, int solver_axis) const {
        return gt[3 * j + solver_axis] == 1 ? n[solver_axis] : n[solver_axis] + bc[solver_axis];
    }
};

// Validate the coupling, compute grid types by two-colouring the coupling
// graph (equivalent to the fixed-point closure of starmap_solver.m:80-98),
// choose kernel axes.  Returns "" on success or an error message.
std::string build_geometry(Geometry &g);

struct TileShape {
    int tile[3] = {1, 1, 1};     // cells per tile per kernel axis
    int lds[3] = {1, 1, 1};      // tile + 1 halo per axis
    int slot_elems = 1;
    int threads = 256;           // cells a tile's staged coefficient fields hold
    bool pair = false;           // two cells per cell thread
    int cells() const { return tile[0] * tile[1] * tile[2]; }
};

// tile of tx*ty*tz <= 256 cells (256 threads, one cell each), or of up to
// 512 cells with two neighbours along the first axis per thread (pair: the
// slot's rows of lds[0] = tx+1 elements give lds[0]/2 threads each, at most
// 256 in all); "" or an error
std::string make_tile(const Geometry &g, int tx, int ty, int tz, TileShape &ts);
TileShape choose_tile(const Geometry &g, int elem_bytes);

struct Plan {
    std::vector<DevGroup> groups;
    std::vector<DevLoad> loads;
    std::vector<TgtRec> targets;
    int n_slots = 0;
    int n_sources = 0;
    int max_terms = 0;
};

// Build the schedule of one substep kind: targets = odd (kind 0) or even
// (kind 1) components.  slot_capacity bounds the LDS slots, group_size the
// targets per group, max_load_elems the source elements a group may stage
// (the kernel's prefetch registers) and max_loads_per_group its load records.
std::string build_plan(const Geometry &g, const TileShape &ts, int kind, int slot_capacity,
                       int group_size, int max_load_elems, int max_loads_per_group, Plan &out);

// ---- discrete adjoint: the LDS-staged update of one launch ---------------
// The adjoint of a launch with targets of one parity P and sources of the
// other Q updates, per target j and point, lambda_j (first adding the
// transposed flux of mu of the launch after it, whose targets are Q) and
// mu_j, and accumulates dJ/dsigma from the forward flux W_j.  Both fluxes
// are sums of coef * (hi - lo) over LDS-staged tiles with the forward
// geometry: W over the state tiles of the sources, the transposed flux over
// their mu tiles with coefficient -c_ij, staged so that mu is zero outside
// the grid and, for a coupling along an extrapolation axis e on which the
// target is cell-centred, zero on the boundary faces of e (a separate tile
// of the same mu per such axis; its forward flux there is identically zero).
// The cell threads form ADJ_PARTS parts, each updating ADJ_TPH of a group's
// up to ADJ_PARTS*ADJ_TPH targets (consecutive in the order of grid types);
// each part sums a point's gradient over its targets of a grid type, and
// the parts' sums are added in a fixed order.
constexpr int ADJ_TERMS = 4;          // terms per kernel axis and list
constexpr int ADJ_TPH = 3;            // targets per part of the cell threads and group
constexpr int ADJ_PARTS = 3;          // parts of the cell threads (256 threads each)
constexpr int ADJ_RECS = ADJ_PARTS * ADJ_TPH;   // records per group
constexpr int ADJ_MAXL = 64;          // source tiles staged per group
constexpr int ADJ_BLOB_LD_OFF = 16;   // load records after the group header
// load records (DevLoad.pad[0]): 0 state tile, 1 mu tile, 2 + e mu tile
// with the boundary faces of kernel axis e zero

struct alignas(16) AdjRec {
    int32_t comp;
    int32_t gq;         // the target's grid type: index among the launch's
    int32_t flush;      // the part's last target of its grid type: store the part's sums
    int32_t jpos;       // comp > 0: dJ/dsigma_s0 accumulates too
    int32_t sig_off, E_off, dE_off, Q_off;   // staged field: element offset of its tile; -1 scalar /
                                             // no source; -2-k: field k of the global table
    int32_t ext[3];
    int32_t pad0;
    int64_t uoff;       // byte offset of the component in a state-layout buffer
    int64_t pad1;
    double sig_s, E_s, dE_s, pad2;
    double F_s[2];      // filter factors of slots k=1,2 (1 without a filter)
    // ADJ_TERMS terms per kernel axis: the forward flux W (state tiles),
    // then the transposed flux (mu tiles); the low sample's offset in the
    // tile (element offset; bytes from the start of LDS on the device) and
    // the coefficient; the high sample lies one stride of the term's axis
    // further
    int32_t off[2 * MAXT];
    double coef[2 * MAXT];
};

struct AdjPlan {
    std::vector<DevGroup> groups; // t0 = index of the group's first record (ADJ_RECS per group)
    std::vector<DevLoad> loads;
    std::vector<AdjRec> targets;  // ADJ_RECS per group, comp = -1: no target
    std::vector<int32_t> gcode;   // grid-type codes of the targets, in order
    int n_slots = 0;
};

// The schedule of the adjoint update for targets odd (kind 0) or even (1).
// Fails (the runtime then uses the gather kernels) if a component couples to
// more than ADJ_TERMS components along an axis or a group does not fit.
// with_w: the forward launches stored W+Q, so neither the state tiles nor
// the forward-flux terms are needed
std::string build_adj_plan(const Geometry &g, const TileShape &ts, int kind, int slot_capacity,
                           bool with_w, AdjPlan &out);

// ---- discrete adjoint: checkpoints, recorded segments, recomputation -----
// The adjoint steps back over segments of `every` steps.  Reversing a
// segment needs the forward's launch outputs ("records") of that segment;
// a segment of m steps has 2m+1 launches, and two consecutive records (one
// per parity) share one state-sized buffer, so its records take m+1
// buffers.  Under a budget of B state-sized buffers the schedule
//   * records the last T segments ("tail") during the forward pass itself:
//     they are never recomputed;
//   * keeps c checkpoint slots for the first H = U - T segments ("head"),
//     reversed from last to first, each recomputed once into one of two
//     alternating record ranges; with fewer slots than head segments the
//     slots are placed by binomial checkpointing (the optimal split of each
//     range of segments, found by dynamic programming over the extra segment
//     advances), which costs further forward advances from a checkpoint.
// T and c minimise the recomputed and re-advanced segments within B.
enum AdjActKind {
    ADJ_FWD_RUN = 0,    // forward pass: segments [s0, s1) in place on the state
    ADJ_FWD_CKPT = 1,   // forward pass: copy the state into buffer dst
    ADJ_FWD_REC = 2,    // forward pass: segment s0 out of place into record range `range`
    ADJ_FWD_FINAL = 3,  // forward pass: the final state from record range `range` into the state
    ADJ_ADVANCE = 4,    // reverse: copy buffer src into dst, then segments [s0, s1) in place on dst
    ADJ_RECOMP = 5,     // reverse: segment s0 from buffer src into record range `range`
    ADJ_REVERSE = 6,    // reverse: adjoint of segment s0 from record range `range`, start state src
};
// buffers other than checkpoint slots (>= 0)
constexpr int ADJ_NONE = -1;
constexpr int ADJ_WORK = -2;        // scratch state of the re-advanced segments
constexpr int ADJ_TAIL_START = -3;  // the state at the start of the tail
constexpr int ADJ_PREV = -4;        // the previous segment's last two records

struct AdjAction {
    int32_t kind, s0, s1, src, dst, range;
};

struct AdjSchedule {
    int32_t segments = 0;       // U
    int32_t seg_steps = 0;      // steps per segment (the last may be shorter)
    int32_t tail = 0;           // T
    int32_t slots = 0;          // c
    int32_t work = 0;           // 1 if the schedule uses the scratch state
    int32_t ranges = 0;         // record ranges of rec_factor*(seg_steps+1) buffers each
    int32_t rec_factor = 1;
    int64_t recomputed = 0;     // head segments recomputed into records
    int64_t extra_advances = 0; // further segment advances from checkpoints
    int64_t buffers = 0;        // state-sized buffers at the peak
    std::vector<AdjAction> fwd, rev;
};

// buffers < 0: no budget (a checkpoint per segment, no tail); rec_factor:
// state-sized buffers per record of two launches (1: the launches' states; 2:
// and the W+Q they stored)
std::string plan_adjoint(int64_t n_steps, int every, int64_t buffers, int rec_factor,
                         AdjSchedule &out);

}  // namespace smp
