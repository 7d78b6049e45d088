// Device kernels of the StaRMAP time-stepping core and their launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "smp_plan.h"

namespace smp {

constexpr int GMAX = 8;       // targets per group
constexpr int MAXL = 19;      // source loads per group (as many as a blob holds)
constexpr int MAX_GROUPS = 256;
constexpr int MAX_CFIELDS = 12; // coefficient fields staged per tile

// Per-group record blob, packed by the host and copied into LDS by the
// kernel: header {loads, targets}, GMAX target records, MAXL load records.
constexpr int BLOB_SZ = 3072;
constexpr int BLOB_REC_OFF = 16;
constexpr int BLOB_LD_OFF = BLOB_REC_OFF + GMAX * (int)sizeof(TgtRec);
constexpr int CPS_MAX = 16;   // DMA chunks per LDS slot
// Waves copying sources into LDS beside cw cell waves (4 or 8).  The copies
// are issued 256 bytes per instruction, so more copy waves keep more bytes
// in flight: 8 beside 8 cell waves (one workgroup of 16 waves per CU), 4
// beside 4 (two workgroups of 8 waves per CU).  The batched kernel has 8
// beside 8 FP32 cell waves with records of up to 10 terms (1% faster on the
// 64-problem sweep than 4), and 4 otherwise: its FP64 variants and those
// with 12-term records need more than the 128 VGPRs a 16-wave workgroup
// leaves.  nt: the record's term count (the template's NT, or the run-time
// count it is chosen by: <= 8, <= 10, more).
constexpr int copy_waves(int cw) { return cw == 8 ? 8 : 4; }
template <typename T>
constexpr int batch_copy_waves(int cw, int nt) { return sizeof(T) == 4 && cw == 8 && nt <= 10 ? 8 : 4; }
constexpr int kthreads(int cw) { return (cw + copy_waves(cw)) * 64; }
template <typename T>
constexpr int batch_threads(int cw, int nt) { return (cw + batch_copy_waves<T>(cw, nt)) * 64; }

template <typename T>
struct SubstepArgs {
    T *U;                     // target values read
    T *out;                   // target values written (U for an update in place)
    const T *src;             // source components (U for an update in place)
    T *sink;                  // 1024 elements written by lanes without a target cell
    int64_t slab;             // elements per component
    int32_t pitch0, ext1;     // storage index = i0 + pitch0*(i1 + ext1*i2)
    int32_t n[3], bc[3];      // per kernel axis
    int32_t tile[3], ntile[3], lds[3];
    int32_t slot_elems;       // lds[0]*lds[1]*lds[2]; lds[0] is a power of two
    int32_t lds0_shift;       // log2(lds[0])
    float inv_lds1;           // 1/lds[1]
    int32_t chunks;           // DMA chunks (64 lanes x 4 bytes) per slot
    int32_t n_tiles;          // total tiles
    int32_t ngroups;
    const unsigned char *blobs;   // ngroups * BLOB_SZ bytes
    const void *const *cfields;   // coefficient fields staged per tile
    int32_t ncf;
    int32_t E[3];             // storage extents per kernel axis
    double half_dt;
    int32_t napp;             // half-step applications with the same flux (1 or 2)
    int32_t fslot[2];         // filter slot of application 0 and 1
    int32_t pair;             // two cells per cell thread (tiles of more than 256 cells)
    T *wout;                  // if not null: W+Q of every target and point (state layout; the
                              // discrete adjoint's record of the launch), without GEN coefficients
};

// One problem's substep in a batched launch: its arguments and the index of
// its first work unit (units unit0 .. unit0 + a.n_tiles - 1 are its tiles).
template <typename T>
struct BatchEntry {
    SubstepArgs<T> a;
    int32_t unit0;
    int32_t pad;
};

// Device-resident descriptor of one batched launch.
template <typename T>
struct BatchLaunch {
    const BatchEntry<T> *entries;
    int32_t n_entries;
    int32_t n_units;
    int *counter;             // work-unit counter, zero before the launch
};

// LDS bytes the kernel needs besides the source slots and the staged
// coefficient fields (ncf * threads elements)
size_t substep_lds_overhead();

// nt: terms per target the kernel unrolls (8 or 12); gen: some coefficient
// is a field read from global memory (more fields than the LDS stages, or a
// spatially varying filter); cw: cell waves (4, or 8 with two waves per
// SIMD on every cell); a.pair: two cells per cell thread (FP32 state)
template <typename T>
hipError_t launch_substep(const SubstepArgs<T> &a, int nt, bool gen, int cw, size_t lds_bytes,
                          hipStream_t s);

// Persistent batched substep launch over `grid` workgroups; bl is a device
// pointer.  nt / gen as for launch_substep, over all entries.
// napp: the half-step applications of every entry of the launch (1 or 2)
// pair: two cells per cell thread (every entry of the launch)
template <typename T>
hipError_t launch_batch(const BatchLaunch<T> *bl, int nt, bool gen, int cw, int napp, bool pair,
                        int grid, size_t lds_bytes, hipStream_t s);

// workgroups of the batched kernel resident per CU with lds_bytes of LDS
template <typename T>
hipError_t batch_occupancy(int nt, bool gen, int cw, int napp, bool pair, size_t lds_bytes,
                           int *blocks);

// p[0..n) = v (a kernel, so that it can be captured in a graph)
hipError_t launch_fill(int *p, int n, int v, hipStream_t s);

template <typename T>
hipError_t launch_decay(const T *sig, T *E, int64_t count, double dt, hipStream_t s);

template <typename T>
hipError_t launch_scale(T *U, const T *F, double F_s, int64_t count, hipStream_t s);

template <typename T>
hipError_t launch_interp(const T *U0, const T *U, double *out, int64_t count, double lambda,
                         hipStream_t s);

// FP64 partial sums of one component: per block (sum, weighted sum of
// squares, non-finite count) into part[3*block]; weights from the staggered
// trapezoid rule per kernel axis.
struct ReduceGeom {
    int32_t ext[3];           // component extent per kernel axis
    int32_t pitch0, ext1;
    int32_t stag[3];          // 1 if the component has n+1 points on this axis
    double cell;              // prod(h)
};

template <typename T>
hipError_t launch_reduce(const T *U, const ReduceGeom &g, double *part, int blocks,
                         hipStream_t s);

hipError_t launch_reduce_final(const double *part, int blocks, double *out3, hipStream_t s);

// Sets *flag to 1 if any of the count elements of U is NaN or infinite
// (the flag is not cleared here).
template <typename T>
hipError_t launch_finite(const T *U, int64_t count, int *flag, hipStream_t s);

// Moment norms of one component: sum of squares and, per kernel axis, of
// squared differences of neighbours in the reference's output order
// (rot[e]: periodic face axis, stored index (f+1) mod n).
struct NormGeom {
    int32_t ext[3];
    int32_t pitch0, ext1;
    int32_t rot[3];
};

// out4 = {sum u^2, sum d0^2, sum d1^2, sum d2^2}; part holds 4*blocks doubles
template <typename T>
hipError_t launch_norm(const T *U, const NormGeom &g, double *part, int blocks, double *out4,
                       hipStream_t s);

// Linear interpolation of one component at a point: up to 8 slab offsets
// and weights (nan: the point is outside the component's grid).
struct ProbeComp {
    int64_t off[8];
    double w[8];
    int32_t n, nan;
};

template <typename T>
hipError_t launch_probe(const T *U, int64_t slab, const ProbeComp *pc, int n, double *out,
                        hipStream_t s);

}  // namespace smp

namespace smp {

// ---- discrete adjoint of the time step ------------------------------------
// One coupling term of the adjoint tables: component `k` with coefficient
// -M(j,k)/h (the forward kernel's coefficient).  A component's terms are
// listed by kernel axis (the forward's order), each active axis's list
// holding at least one group of ADJ_GROUP terms, padded with coef = 0 terms.
// Every component coupled to another along an axis has the opposite
// staggering on that axis, so the sample offsets follow from the axis and
// the grid type of the component that owns the list.
struct AdjTerm {
    int32_t k;
    int32_t pad;
    double coef;
};
constexpr int ADJ_GROUP = 4;
// rows of points per thread in the transposed-flux kernel
constexpr int ADJ_ROWS = 2;

// Per component: its coefficients in the forward half-step, its extents and
// grid types per kernel axis and its ranges in the term tables (terms by
// target row and by source column).
template <typename T>
struct AdjComp {
    const T *sig;             // sigma field (slab layout) or null: sig_s
    const T *E;               // expm1div(-sigma dt/2) field or null: E_s
    const T *dE;              // dE/dsigma field or null: dE_s
    const T *Q;               // source field or null (no source)
    double sig_s, E_s, dE_s;
    double F[2];              // filter factors of slots k=1,2 (1 without filter)
    int32_t ext[3];
    int32_t cell[3];          // 1 if cell-centred on the kernel axis
    int32_t ra[4];            // forward terms on kernel axis e: [ra[e], ra[e+1])
    int32_t ca[4];            // transposed terms (this component as source)
};

template <typename T>
struct AdjArgs {
    const AdjComp<T> *comp;
    const AdjTerm *rows;      // terms by target: k = source component
    const AdjTerm *cols;      // terms by source: k = target component
    const int32_t *list;      // components of the parity this launch updates,
    int32_t n_list;           // grouped by grid type
    int32_t n_grp;            // grid types in the list
    int32_t grp0[9];          // list range of each grid type
    int32_t gcode[8];         // grid type code (gtx-1)*4+(gty-1)*2+(gtz-1)
    const T *U;               // forward state during the substep (sources)
    const T *U0;              // targets before the substep (rows kernel)
    const T *W;               // W+Q of the targets as the forward launch stored it, or null
                              // (the update kernel then computes the flux W itself)
    T *lam;                   // adjoint state
    T *mu;                    // adjoint of the flux W of the targets
    double *Ga, *Gs;          // dJ/dsigma_a, dJ/dsigma_s0 per grid type code
                              // and point (8 slabs each, FP64)
    int64_t slab;
    int32_t pitch0, ext1;
    int32_t E[3];             // storage extents per kernel axis
    int32_t n[3], bc[3];      // cells and boundary per kernel axis
    double half_dt;
    int32_t napp;
    int32_t fslot[2];
    int32_t ndim;             // active kernel axes (1..3)
};

// adjoint of the half-step applications of one substep for the targets in
// a.list: mu = B*lambda', lambda = A*lambda', and per grid type the sum over
// its targets of lambda'*(dA u + dB (W+Q)) added to Ga (all targets) and Gs
// (targets j > 1): one thread per point and grid type
// dE[i] = dE/dsigma of expm1div(-sigma*dt/2) at sig[i]
template <typename T>
hipError_t launch_dedsig(const T *sig, T *dE, int64_t count, double dt, hipStream_t s);

// fuse: first lambda += the transposed flux of mu (the other parity's
// targets of the launch after this one), as launch_adj_transpose for the
// sources of that launch
template <typename T>
hipError_t launch_adj_update(const AdjArgs<T> &a, bool fuse, hipStream_t s);

// lambda_k += sum_j coef_jk (D^T mu_j), k in a.list (the substep's sources)
template <typename T>
hipError_t launch_adj_transpose(const AdjArgs<T> &a, hipStream_t s);

// dst[i] = src[i] for a contiguous range of elements
template <typename T>
hipError_t launch_copy(T *dst, const T *src, int64_t count, hipStream_t s);

// ---- the LDS-staged adjoint update (one launch's targets) -----------------
// Arguments of adj_lds_kernel: the tile decomposition and LDS layout of the
// substep kernel (one cell per cell thread, 4 cell waves beside 4 copy
// waves), the plan of one parity (build_adj_plan), and per launch the
// forward states and napp/filter slots of the launch it reverses.
template <typename T>
struct AdjTileArgs {
    const T *U;               // the launch's sources (the state after launch L-1)
    const T *U0;              // its targets before it (after launch L-2)
    const T *W;               // W+Q its targets had, as the forward launch stored it (or null)
    T *lam, *mu;              // adjoint state and adjoint of the flux (state layout)
    double *Ga, *Gs;          // dJ/dsigma_a, dJ/dsigma_s0: 8 grid-type slabs each
    T *sink;                  // 1024 elements written by lanes without a target point
    int64_t slab;
    int32_t pitch0, ext1;
    int32_t n[3], bc[3];
    int32_t tile[3], ntile[3], lds[3];
    int32_t slot_elems, lds0_shift;
    float inv_lds1;
    int32_t chunks, n_tiles, ngroups;
    const unsigned char *blobs;   // per group: header {loads, targets} and load records
    const AdjRec *recs;           // ADJ_RECS records per group
    const unsigned *zeros;        // 64 zero dwords (boundary faces of masked mu tiles)
    const void *const *cfields;   // coefficient fields staged per tile
    int32_t ncf;
    const void *const *gfields;   // coefficient fields read from global memory (off -2-k: field k)
    int32_t E[3];
    double half_dt;
    int32_t napp, fslot[2];
    int32_t ngt;                  // grid types of the targets, their codes and extents
    int32_t gcode[4];
    int32_t gext[4][3];
};

// LDS bytes before the source slots: blob ring, staged fields, gradient
// accumulators of ngt grid types
size_t adj_lds_overhead(int ncf, int ngt, int elem);
constexpr int ADJ_THREADS = 1024;
// the tile shape the kernel is compiled for, per number of kernel axes
// (cells per axis; the LDS slot adds a one-cell halo on each axis)
constexpr int adj_tile_dim(int ndim, int e) {
    return ndim == 1 ? (e == 0 ? 255 : 1) : ndim == 2 ? (e == 0 ? 63 : e == 1 ? 4 : 1)
                                                       : (e == 0 ? 31 : e == 1 ? 4 : 2);
}

// ndim: active kernel axes; fuse: add the transposed flux of the launch
// after first (mu tiles staged); gen: some coefficient is a field read from
// global memory (more fields than the LDS stages); a.W not null: W+Q from
// the forward launch's record (the plan then stages no state tiles)
template <typename T>
hipError_t launch_adj_lds(const AdjTileArgs<T> &a, int ndim, bool fuse, bool gen, size_t lds_bytes,
                          hipStream_t s);

}  // namespace smp
