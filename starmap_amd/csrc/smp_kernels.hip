// gfx950 kernels of the StaRMAP time-stepping core.
//
//  substep_kernel  (a)+(b)+(c): one Strang substep of starmap_solver.m:313-359
//                  or :361-414 for every target component of one parity,
//                  with the exact exponential decay factor and the substep
//                  filter fused into the update.
//  decay_kernel    (b): E = expm1div(-sigma*dt/2), starmap_solver.m:267-274,
//                  :531-535, for spatially varying sigma.
//  scale_kernel    (c): full-step filter U{j} .*= F, starmap_solver.m:417-421.
//  interp_kernel   (d): (1-lambda)*U0 + lambda*U, starmap_solver.m:425-428.
//  reduce_kernel   (d): FP64 sums / weighted squares for observables.
#include "smp_kernels.h"

#include <cmath>

namespace smp {

namespace {

__device__ __forceinline__ int wrap_or_clamp(int i, int n, int bc, int ext) {
    if (bc == 0) {
        // periodic: every grid of the axis has n points
        if (i < 0) i += n;
        if (i >= n) i %= n;
        return i;
    }
    // extrapolation: clamping reproduces the extension vectors
    // [1, 1:n, n] / [1:n+1] of starmap_solver.m:127-132 (zero differences on
    // the boundary faces)
    return i < 0 ? 0 : (i >= ext ? ext - 1 : i);
}

// Blocks are dealt round-robin to the 8 XCDs of the MI355X; give each XCD a
// contiguous run of tiles so that neighbouring tiles (which share halo
// rows/columns) are resident in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

}  // namespace

static_assert(BLOB_LD_OFF + MAXL * (int)sizeof(DevLoad) <= BLOB_SZ, "blob layout");
static_assert(sizeof(DevLoad) == 32 && sizeof(TgtRec) % 16 == 0, "record layout");

// LDS: a ring of 3 group blobs | staged coefficient fields | source slots
size_t substep_lds_overhead() { return 3 * (size_t)BLOB_SZ; }

typedef __attribute__((address_space(3))) void *lds_ptr_t;

// Pointers into device memory taken from argument structs that live in
// memory (the batched kernel) are generic pointers to the compiler, and
// generic accesses compile to flat instructions with conservative waits;
// these casts state that they address global memory.
#define SMP_GLOBAL __attribute__((address_space(1)))
#define SMP_CONST __attribute__((address_space(4)))
#ifndef SMP_SCALAR_RECORDS
#define SMP_SCALAR_RECORDS 1
#endif
template <typename P>
__device__ __forceinline__ SMP_GLOBAL P *glb(P *p) {
    return (SMP_GLOBAL P *)p;
}

// One workgroup = one spatial tile of up to 256 cells.  The first CW waves
// (cell waves) own one cell per thread and update the targets; the last NCW
// waves copy: source tiles and the group records go from HBM straight into
// LDS (global_load_lds, no registers).  While the cell waves update the
// targets of group g from LDS, the copy waves bring in the source tiles of
// group g+1 and the records of group g+2, and the cell waves load the
// target values of group g+1; one barrier per group.  Only the copy waves
// have copies in flight, so the cell waves never wait on them before reading
// LDS.  The schedule never lets group g+1 copy into a slot that group g
// reads.
//
// NAPP: the number of half-step applications, compiled in (the batched
// kernel has one launch per application count).
// AT: SubstepArgs<T> (kernel argument) or the same in the constant address
// space (an entry of the batched kernel, read with scalar loads).
// CW cell waves (4 or 8): with 8, two waves per SIMD share each cell, each
// updating half of a group's targets.
//
// PAIR: every cell thread updates two cells, neighbours along the first
// kernel axis (tiles of up to 512 cells; with 8 cell waves, two threads share
// each pair of cells, as they share a cell in the one-cell path).  A term's samples of
// both cells are adjacent in LDS, so each term costs one address computation
// and two LDS reads of neighbouring elements for the two cells, and a target
// record read serves both; each cell's arithmetic is that of the one-cell
// path, so the results agree bit for bit.
// WOUT: also store W+Q of every target and point to a.wout (for the discrete
// adjoint, which then needs no flux of its own for the gradient).
template <typename T, int NT, int NAPP, bool GEN, int CW, int NCW, bool WARM, bool PAIR, bool WOUT,
          typename AT>
__device__ __forceinline__ void substep_tile(const AT &a, const int tile, unsigned char *smem_raw) {
    constexpr int NCT = CW * 64;              // cell threads
    constexpr int TPT = GMAX * 4 / CW;        // targets of a group per cell thread
    constexpr int PC = PAIR ? 2 : 1;          // cells per cell thread
    unsigned char *ring = smem_raw;
    T *cfa = reinterpret_cast<T *>(smem_raw + 3 * BLOB_SZ);
    T *lds = cfa + a.ncf * 256 * PC;
    const int tid = threadIdx.x;
    // arguments the update reads for every target, held in registers (a
    // scalar reload inside the loop would also wait for the LDS reads in
    // flight, which share its counter)
    const double half_dt = __builtin_bit_cast(
        double, ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                     (int)(__builtin_bit_cast(uint64_t, (double)a.half_dt) >> 32))
                 << 32) |
                    (uint32_t)__builtin_amdgcn_readfirstlane(
                        (int)(uint32_t)__builtin_bit_cast(uint64_t, (double)a.half_dt)));
    const float half_dt_f = (float)half_dt;
    const int fslot0 = __builtin_amdgcn_readfirstlane(a.fslot[0]);
    const int fslot1 = __builtin_amdgcn_readfirstlane(a.fslot[1]);
    const int lane = tid & 63;
    // wave-uniform (waves are 64 threads, NCT a multiple of 64): copy waves
    // and cell waves run separate loops with the same number of barriers
    const bool copier = __builtin_amdgcn_readfirstlane(tid) >= NCT;
    const int ctid = tid & 255;               // the cell of a cell thread
    const int half = CW == 8 ? (tid >> 8) & 1 : 0;
    int b = tile;
    const int t0 = b % a.ntile[0];
    b /= a.ntile[0];
    const int t1 = b % a.ntile[1];
    const int t2 = b / a.ntile[1];
    const int o0 = t0 * a.tile[0], o1 = t1 * a.tile[1], o2 = t2 * a.tile[2];
    // The cell of a cell thread.  Where lds[0]*tile[1]*tile[2] <= 256 (every
    // tile shape the library picks: lds[0] = tile[0] + 1), thread ctid takes
    // the slot column ctid mod lds[0], so a wave's LDS reads are consecutive
    // elements; numbered along the tile instead, the threads would skip the
    // slot's halo column once per tile row, and two threads of one read pass
    // would meet in the same bank.  A thread on a column or row without a
    // cell reads the address of the nearest cell (a broadcast, in the slot).
    const bool by_slot = a.lds[0] * a.tile[1] * a.tile[2] <= 256;
    int c0, c1, c2;
    bool has_cell, has_cell1 = false;
    if (PAIR) {
        // slot columns 2p and 2p+1 of row r (the host checks that
        // lds[0]/2 * tile[1] * tile[2] <= 256); column lds[0]-1 is the halo
        const int r = ctid >> (a.lds0_shift - 1);
        c0 = (ctid & ((a.lds[0] >> 1) - 1)) << 1;
        c1 = r % a.tile[1];
        c2 = r / a.tile[1];
        has_cell = tid < NCT && c0 < a.tile[0] && c2 < a.tile[2];
        has_cell1 = has_cell && c0 + 1 < a.tile[0];
    } else if (by_slot) {
        const int r = ctid >> a.lds0_shift;
        c0 = ctid & (a.lds[0] - 1);
        c1 = r % a.tile[1];
        c2 = r / a.tile[1];
        has_cell = tid < NCT && c0 < a.tile[0] && c2 < a.tile[2];
    } else {
        c0 = ctid % a.tile[0];
        c1 = (ctid / a.tile[0]) % a.tile[1];
        c2 = ctid / (a.tile[0] * a.tile[1]);
        has_cell = tid < NCT && ctid < a.tile[0] * a.tile[1] * a.tile[2];
    }
    const int g0 = o0 + c0, g1 = o1 + c1, g2 = o2 + c2;
    const bool in_store = has_cell && g0 < a.E[0] && g1 < a.E[1] && g2 < a.E[2];
    const bool in_store1 = has_cell1 && g0 + 1 < a.E[0] && g1 < a.E[1] && g2 < a.E[2];
    const int lbase = min(c0, a.tile[0] - 1) + a.lds[0] * (min(c1, a.tile[1] - 1) + a.lds[1] * min(c2, a.tile[2] - 1));
    const int64_t gidx = g0 + (int64_t)a.pitch0 * (g1 + (int64_t)a.ext1 * g2);
    const int S = a.slot_elems;
    const int ng = a.ngroups;
    constexpr int DPE = sizeof(T) / 4;      // dwords per element
    constexpr int EPC = 64 / DPE;           // elements per 64-lane DMA chunk

    auto hdr_nl = [&](int buf) { return reinterpret_cast<const int *>(ring + buf * BLOB_SZ)[0]; };
#if !SMP_SCALAR_RECORDS
    auto recs = [&](int buf) {
        return reinterpret_cast<const TgtRec *>(ring + buf * BLOB_SZ + BLOB_REC_OFF);
    };
#endif
    auto ldrec = [&](int buf) {
        return reinterpret_cast<const DevLoad *>(ring + buf * BLOB_SZ + BLOB_LD_OFF);
    };

    // ---- copy wave -------------------------------------------------------
    // Per tile, the dword offset in a source slab of the element each lane
    // copies in chunk cc for a face-centred source (no origin shift), with
    // boundary wrapping / clamping resolved (-1 past the slot), and per axis
    // a bit per chunk for lanes at coordinate 0.  A cell-centred source looks
    // one cell lower on its shifted axes: its coordinate is x-1, or at x == 0
    // n-1 (periodic) / 0 (extrapolation, clamped).  Where no lane is at
    // x == 0, a uniform shift of the slab base does it.
    int po[CPS_MAX];
    int zb[3] = {0, 0, 0};
    const int st[3] = {DPE, a.pitch0 * DPE, a.pitch0 * a.ext1 * DPE};
    int wd[3];           // offset change at x == 0 for a shifted axis
#pragma unroll
    for (int e = 0; e < 3; ++e) wd[e] = a.bc[e] == 0 ? (a.n[e] - 1) * st[e] : 0;
    bool low_inner = false;
    if (copier) {
#pragma unroll
        for (int cc = 0; cc < CPS_MAX; ++cc) {
            const int e = cc * EPC + lane / DPE;
            po[cc] = -1;
            if (cc < a.chunks && e < S) {
                const int l0 = e & (a.lds[0] - 1);
                const int row = e >> a.lds0_shift;
                const int l2 = (int)(((float)row + 0.5f) * a.inv_lds1);   // exact for rows < 2^16
                const int l1 = row - l2 * a.lds[1];
                const int x0 = wrap_or_clamp(o0 + l0, a.n[0], a.bc[0], a.n[0] + a.bc[0]);
                const int x1 = wrap_or_clamp(o1 + l1, a.n[1], a.bc[1], a.n[1] + a.bc[1]);
                const int x2 = wrap_or_clamp(o2 + l2, a.n[2], a.bc[2], a.n[2] + a.bc[2]);
                po[cc] = (x0 + a.pitch0 * (x1 + a.ext1 * x2)) * DPE + (DPE == 2 ? (lane & 1) : 0);
                zb[0] |= (x0 == 0) << cc;
                zb[1] |= (x1 == 0) << cc;
                zb[2] |= (x2 == 0) << cc;
            }
        }
        // x == 0 occurs at the low boundary and, periodic, where the tile
        // wraps at the top; a wave without such a lane shifts uniformly
        int zany = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e)
            if (a.lds[e] > 1) zany |= zb[e];
        low_inner = !__any(zany != 0);
    }
    auto blob_dma = [&](int gq, int buf) {
        // three 1 KB chunks, 16 bytes a lane, by the first copy wave
        if (tid >= NCT + 64) return;
#pragma unroll
        for (int k = 0; k < BLOB_SZ / 1024; ++k) {
            SMP_GLOBAL const unsigned char *src = glb(a.blobs) + (size_t)gq * BLOB_SZ + k * 1024 + lane * 16;
            __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(ring + buf * BLOB_SZ + k * 1024), 16, 0,
                                             0);
        }
    };
    // the source copies of a group whose nl load records are at ld (in the
    // LDS ring, or in the plan in global memory)
    auto src_dma_recs = [&](const int nl, const auto *ld) {
        // one load record per lane, read before the first copy is issued (an
        // LDS read behind an outstanding copy would have to wait for it):
        // comp | slot << 20 | (origin-shift bits) << 28
        const int cw = (tid >> 6) - CW;  // copy wave cw: every NCW-th load
        int myrec = 0;
        if (lane < nl) {
            const auto &d = ld[lane];
            myrec = d.comp | (d.slot << 20) |
                    (((d.org[0] & 1) | ((d.org[1] & 1) << 1) | ((d.org[2] & 1) << 2)) << 28);
        }
        for (int li = cw; li < nl; li += NCW) {
            const int rec = __builtin_amdgcn_readlane(myrec, li);
            const int comp = rec & 0xFFFFF;
            const int slot = (rec >> 20) & 0xFF;
            const int m = (rec >> 28) & 7;   // bit e: cell-centred on kernel axis e
            SMP_GLOBAL const unsigned *cb = glb(reinterpret_cast<const unsigned *>(a.src + (int64_t)comp * a.slab));
            unsigned char *dst = reinterpret_cast<unsigned char *>(lds + slot * S);
            if (low_inner) {
                // every shifted coordinate is x-1: one uniform base shift
                SMP_GLOBAL const unsigned *cbs = cb - ((m & 1) ? st[0] : 0) - ((m & 2) ? st[1] : 0) -
                                                 ((m & 4) ? st[2] : 0);
#pragma unroll
                for (int cc = 0; cc < CPS_MAX; ++cc) {
                    // lanes past the end of the slot copy nothing
                    if (cc < a.chunks && po[cc] >= 0)
                        __builtin_amdgcn_global_load_lds(cbs + po[cc], (lds_ptr_t)(dst + cc * 256),
                                                         4, 0, 0);
                }
            } else {
#pragma unroll
                for (int cc = 0; cc < CPS_MAX; ++cc) {
                    if (cc < a.chunks && po[cc] >= 0) {
                        int off = po[cc];
#pragma unroll
                        for (int e = 0; e < 3; ++e)
                            if (m & (1 << e)) off += ((zb[e] >> cc) & 1) ? wd[e] : -st[e];
                        __builtin_amdgcn_global_load_lds(cb + off, (lds_ptr_t)(dst + cc * 256), 4, 0,
                                                         0);
                    }
                }
            }
        }
    };

    auto src_dma = [&](int buf) { src_dma_recs(__builtin_amdgcn_readfirstlane(hdr_nl(buf)), ldrec(buf)); };
    // group gq's copies from its load records in the plan itself: the first
    // group of a tile needs no LDS blob, so its copies start in the
    // prologue's first phase
    auto src_dma_global = [&](int gq) {
        const SMP_CONST unsigned char *gb = (const SMP_CONST unsigned char *)a.blobs + (int64_t)gq * BLOB_SZ;
        src_dma_recs(*(const SMP_CONST int *)gb,
                     reinterpret_cast<SMP_GLOBAL const DevLoad *>(glb(a.blobs) + (int64_t)gq * BLOB_SZ +
                                                                  BLOB_LD_OFF));
    };

    // ---- cell waves ------------------------------------------------------
    // Target loads and stores are issued by every lane (lanes outside a
    // target's extent read element 0 and write to the sink), so that the
    // number of memory operations per group is fixed and the compiler can
    // wait for a prefetched target value with an exact count instead of
    // draining every load and store in flight.
    SMP_GLOBAL const char *ug_ld = glb(reinterpret_cast<const char *>(a.U)) + (in_store ? gidx : 0) * sizeof(T);
    SMP_GLOBAL char *ug_st = glb(reinterpret_cast<char *>(a.out)) + gidx * sizeof(T);
    // the second cell of a pair reads its neighbour's address when it is
    // outside the storage (the value is not used)
    const int d1 = in_store1 ? (int)sizeof(T) : 0;
    SMP_GLOBAL T *sink = glb(a.sink);
    SMP_GLOBAL char *wo_st = WOUT ? glb(reinterpret_cast<char *>(a.wout)) + gidx * sizeof(T) : nullptr;
    // The cell waves' records of group gq: with SMP_SCALAR_RECORDS, read
    // with scalar loads from the plan in global memory (every workgroup
    // reads the same records, which stay in the scalar cache and L2);
    // otherwise from the group's blob in the LDS ring.
#if SMP_SCALAR_RECORDS
    const int half_u = __builtin_amdgcn_readfirstlane(half);
    auto rec_of = [&](int gq, int buf) {
        (void)buf;
        return (const SMP_CONST TgtRec *)(a.blobs + (int64_t)gq * BLOB_SZ + BLOB_REC_OFF) + half_u * TPT;
    };
#else
    auto rec_of = [&](int gq, int buf) {
        (void)gq;
        return recs(buf) + half * TPT;
    };
#endif
    auto load_targets = [&](int gq, int buf, T *u) {
        const auto r = rec_of(gq, buf);
#pragma unroll
        for (int i = 0; i < TPT; ++i) {
            u[PC * i] = *reinterpret_cast<SMP_GLOBAL const T *>(ug_ld + r[i].uoff);
            if (PAIR) u[PC * i + 1] = *reinterpret_cast<SMP_GLOBAL const T *>(ug_ld + r[i].uoff + d1);
        }
    };
    // byte address of this thread's cell in the first source slot; term
    // offsets are byte offsets from the start of LDS
    const int lb = lbase * (int)sizeof(T);
    // The flux sum W computes in the state precision (FP32 terms with FP32
    // coefficients for the FP32 state).  The half-step of the FP64 state,
    // and of the FP32 state with fields read from global memory (GEN), is
    // computed in FP64.  The FP32 state otherwise computes it in FP32 as a
    // change added to u once,
    //     u <- u + (c - (1-F) u - (1-F) c),  c = (dt/2 E)(W+Q) - (dt/2 s E) u,
    // so that every factor near 1 enters through its small part x = 1-F or
    // dt/2 s E: a factor rounded to FP32 once and applied every half-step
    // biases the state by the same relative amount at each application, an
    // error that grows linearly with the step count (about 1e-5 after 500
    // filtered steps if F itself is rounded), where the rounded x leaves a
    // bias of x times that.  Applying decay, source and filter one after the
    // other, each rounded, lost the changes below half an ulp of u at each of
    // three roundings: 6.5 times the FP64-update error after the 915 steps of
    // the 2D P39 lattice.
    typedef T A;
    typedef double D;
    // coefficient: staged field (off >= 0), scalar (off == -1) or, in the
    // GEN variant only, a field read from global memory (off == -2)
    auto coef = [&](int off, const void *p, double s, int k) -> D {
        const D staged = (D)cfa[(off > 0 ? off : 0) + PC * ctid + k];
        D v = off >= 0 ? staged : s;
        if (GEN && off == -2)
            v = (k ? in_store1 : in_store) ? (D)glb(reinterpret_cast<const T *>(p))[gidx + k] : 0.0;
        return v;
    };
    // Every group holds GMAX target records; unused ones are zero records
    // (extent 0, terms in the zero tile), so the loop has no branches and the
    // compiler can overlap one target's LDS reads with another's arithmetic.
    auto update = [&](int gq, int buf, const T *u_in) {
        const auto rb = rec_of(gq, buf);
#pragma unroll
        for (int i = 0; i < TPT; ++i) {
            const auto &r = rb[i];
            const bool act = has_cell & (g0 < r.ext[0]) & (g1 < r.ext[1]) & (g2 < r.ext[2]);
            const bool act1 = has_cell1 & (g0 + 1 < r.ext[0]) & (g1 < r.ext[1]) & (g2 < r.ext[2]);
            // W = -sumcell(...): the coefficients carry -M(j,k)/h; padding
            // terms add coef 0 times a difference within the zero tile
            A w[PC];
#pragma unroll
            for (int k = 0; k < PC; ++k) w[k] = 0;
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const int toff = r.t[q].off, tstride = r.t[q].stride;
                const unsigned char *lo = smem_raw + (toff + lb);
                const A c = sizeof(T) == 4 ? (A)r.t[q].coef32[0] : (A)r.t[q].coef;
#pragma unroll
                for (int k = 0; k < PC; ++k)
                    w[k] += c * ((A)*reinterpret_cast<const T *>(lo + tstride + k * sizeof(T)) -
                                 (A)*reinterpret_cast<const T *>(lo + k * sizeof(T)));
            }
#pragma unroll
            for (int k = 0; k < PC; ++k) {
                const bool ak = k ? act1 : act;
                if constexpr (sizeof(T) == 4 && !GEN) {
                    // every record value is read before the selects: a read
                    // in one arm only would become a branch around a scalar
                    // load, whose wait also drains the LDS reads in flight
                    const float eps_s = r.f32[0], B_s = r.f32[1], dF0 = r.f32[2], dF1 = r.f32[3],
                                sig_s = r.f32[4];
                    const float sgf = (float)cfa[(r.sig_off > 0 ? r.sig_off : 0) + PC * ctid + k];
                    const float Ef = (float)cfa[(r.E_off > 0 ? r.E_off : 0) + PC * ctid + k];
                    const float Qf = (float)cfa[(r.Q_off > 0 ? r.Q_off : 0) + PC * ctid + k];
                    const float Bf = half_dt_f * Ef;
                    const float B = r.E_off >= 0 ? Bf : B_s;
                    // sigma times dt/2 E: both scalars (the record's product),
                    // or a field times a scalar or a field
                    const float eps = r.sig_off >= 0 ? B * sgf : (r.E_off >= 0 ? Bf * sig_s : eps_s);
                    const float wq = (float)w[k] + (r.Q_off >= 0 ? Qf : 0.0f);
                    if (WOUT) {
                        SMP_GLOBAL T *dw = ak ? reinterpret_cast<SMP_GLOBAL T *>(wo_st + r.uoff + k * sizeof(T))
                                              : sink + PC * tid + k;
                        *dw = (T)wq;
                    }
                    float u = (float)u_in[PC * i + k];
#pragma unroll
                    for (int app = 0; app < NAPP; ++app) {
                        // (u + c)(1 - dF) = u + (c - dF u - dF c), with the
                        // change c = dt/2 E (W+Q) - dt/2 s E u: the change is
                        // summed in FP32 relative to its own size and added
                        // to u once, one rounding per half-step
                        const bool s1 = (app == 0 ? fslot0 : fslot1) != 0;
                        const float dF = s1 ? dF1 : dF0;   // 0 without a filter
                        const float c = __builtin_fmaf(B, wq, -(eps * u));
                        float d = __builtin_fmaf(-dF, u, c);
                        d = __builtin_fmaf(-dF, c, d);
                        u = u + d;
                    }
                    SMP_GLOBAL T *dst = ak ? reinterpret_cast<SMP_GLOBAL T *>(ug_st + r.uoff + k * sizeof(T))
                                           : sink + PC * tid + k;
                    *dst = (T)u;
                    continue;
                }
                const D sig = coef(r.sig_off, r.sig_p, r.sig_s, k);
                const D E = coef(r.E_off, r.E_p, r.E_s, k);
                const D wq = (D)w[k] + coef(r.Q_off, r.Q_p, 0.0, k);
                if (WOUT) {
                    SMP_GLOBAL T *dw = ak ? reinterpret_cast<SMP_GLOBAL T *>(wo_st + r.uoff + k * sizeof(T))
                                          : sink + PC * tid + k;
                    *dw = (T)wq;
                }
                D u = (D)u_in[PC * i + k];
#pragma unroll
                for (int app = 0; app < NAPP; ++app) {
                    // U + dt/2*(W+Q - s.*U).*E   (starmap_solver.m:352-354)
                    u = u + (half_dt * (wq - sig * u)) * E;
                    const bool s1 = (app == 0 ? fslot0 : fslot1) != 0;
                    D f = s1 ? r.F_s[1] : r.F_s[0];   // 1 for targets without a filter
                    if (GEN && r.filt) {
                        const void *fp = s1 ? r.F[1] : r.F[0];
                        if (fp && ak) f = (D)glb(reinterpret_cast<const T *>(fp))[gidx + k];
                    }
                    // the state after every half-step is a value of the state
                    // precision, as if stored: fused and separate launches agree
                    u = (D)(T)(u * f);
                }
                SMP_GLOBAL T *dst = ak ? reinterpret_cast<SMP_GLOBAL T *>(ug_st + r.uoff + k * sizeof(T))
                                       : sink + PC * tid + k;
                *dst = (T)u;
            }
        }
    };

    // The cell waves read their target records with scalar loads, and a
    // wait for a scalar load also waits for their LDS reads in flight; a
    // tile's records are far larger than the scalar cache.  With WARM (the
    // batched kernel, whose small tiles change problem and plan often), the
    // last copy wave reads one dword of every 64-byte line of a group's
    // records two phases before the cell waves use them, so that their loads
    // hit.  The plain kernel's tuned schedules hide the misses, and the
    // extra loads cost them 1-2%.
    const bool warmer = WARM && SMP_SCALAR_RECORDS &&
                        __builtin_amdgcn_readfirstlane(tid >> 6) == CW + NCW - 1;
    auto warm_records = [&](int gq) {
        const SMP_CONST int *p = (const SMP_CONST int *)(a.blobs + (int64_t)gq * BLOB_SZ);
        int acc = 0;
#pragma unroll
        for (int k = 0; k < (BLOB_REC_OFF + GMAX * (int)sizeof(TgtRec) + 63) / 64; ++k) acc += p[k * 16];
        asm volatile("" : : "s"(acc));
    };

    // ---- copy waves ----------------------------------------------------------
    // prologue (one barrier: group 0's sources from its records in global
    // memory, beside the records of groups 0 and 1), then one barrier per
    // group: the sources of group g+1 and the records of group g+2 while the
    // cells update group g
    if (copier) {
        if (ng > 0) blob_dma(0, 0);
        if (ng > 1) blob_dma(1, 1);
        if (warmer && ng > 0) warm_records(0);
        if (warmer && ng > 1) warm_records(1);
#if SMP_SCALAR_RECORDS
        if (ng > 0) src_dma_global(0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
#else
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (ng > 0) src_dma(0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
#endif
        for (int gq = 0; gq < ng; ++gq) {
            if (gq + 1 < ng) src_dma((gq + 1) % 3);
            if (gq + 2 < ng) blob_dma(gq + 2, (gq + 2) % 3);
            if (warmer && gq + 2 < ng) warm_records(gq + 2);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        return;
    }

    // ---- cell waves ----------------------------------------------------------
    // the same barriers; after updating group g, the target values of group
    // g+1 are loaded into the other of two register sets (the loop is unrolled
    // by two, so no values move between sets): they arrive while the copy
    // waves finish, and the next phase's first read of them is the only wait
    // for memory in the cell waves
    T uA[TPT * PC], uB[TPT * PC];
    if (half == 0) {
        // every field load in flight before the first LDS write
        T cv[MAX_CFIELDS * PC];
#pragma unroll
        for (int c = 0; c < MAX_CFIELDS; ++c)
            if (c < a.ncf) {
                const T *fc = reinterpret_cast<const T *>(a.cfields[c]);
                cv[PC * c] = glb(fc)[in_store ? gidx : 0];
                if (PAIR) cv[PC * c + 1] = glb(fc)[in_store1 ? gidx + 1 : 0];
            }
#pragma unroll
        for (int c = 0; c < MAX_CFIELDS; ++c)
            if (c < a.ncf) {
                cfa[c * 256 * PC + PC * ctid] = in_store ? cv[PC * c] : (T)0;
                if (PAIR) cfa[c * 256 * PC + PC * ctid + 1] = in_store1 ? cv[PC * c + 1] : (T)0;
            }
    }
    for (int e = tid; e < S; e += NCT) lds[e] = (T)0;   // slot 0: the zero tile
#if SMP_SCALAR_RECORDS
    // the records come from global memory: no barrier before the first loads
    if (ng > 0) load_targets(0, 0, uA);
    __syncthreads();
#else
    __syncthreads();
    if (ng > 0) load_targets(0, 0, uA);
    __syncthreads();
#endif
    for (int gq = 0; gq < ng; gq += 2) {
        update(gq, gq % 3, uA);
        if (gq + 1 < ng) load_targets(gq + 1, (gq + 1) % 3, uB);
        __syncthreads();
        if (gq + 1 >= ng) break;
        update(gq + 1, (gq + 1) % 3, uB);
        if (gq + 2 < ng) load_targets(gq + 2, (gq + 2) % 3, uA);
        __syncthreads();
    }
}

template <typename T, int NT, int NAPP, bool GEN, int CW, bool PAIR, bool WOUT>
__global__ __launch_bounds__(kthreads(CW)) void substep_kernel(const SubstepArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.n_tiles) return;
    substep_tile<T, NT, NAPP, GEN, CW, copy_waves(CW), false, PAIR, WOUT>(a, tile, smem_raw);
}

// Batched persistent kernel: one launch runs a substep of many independent
// problems.  Work units are (entry, tile) pairs, entries ordered by cost per
// tile, heaviest first; every workgroup takes the next unit from a counter
// until none is left, so uneven problems balance over the CUs.
template <typename T, int NT, int NAPP, bool GEN, int CW, bool PAIR>
__global__ __launch_bounds__(batch_threads<T>(CW, NT)) void batch_kernel(const BatchLaunch<T> *bl) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    const int n_units = bl->n_units;
    const int n_entries = bl->n_entries;
    // the entries are not written during the launch: scalar loads
    const SMP_CONST BatchEntry<T> *ent = (const SMP_CONST BatchEntry<T> *)bl->entries;
    for (;;) {
        if (threadIdx.x == 0)
            s_unit = __hip_atomic_fetch_add(glb(bl->counter), 1, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= n_units) break;
        // entry holding unit u: the last entry whose first unit is <= u
        int lo = 0, hi = n_entries - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ent[mid].unit0 <= u)
                lo = mid;
            else
                hi = mid - 1;
        }
        // the entry's arguments, copied into registers as the plain
        // kernel's arguments are (reading them from memory inside the loops
        // costs the update its overlap of LDS reads)
        SubstepArgs<T> a;
        static_assert(sizeof(SubstepArgs<T>) % 4 == 0, "argument dwords");
        const SMP_CONST int *src = (const SMP_CONST int *)&ent[lo].a;
        int *dst = reinterpret_cast<int *>(&a);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(SubstepArgs<T>) / 4); ++k)
            dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
        // the launch's application count is compiled in (a runtime count
        // costs the update its instruction schedule, and both tile paths in
        // one kernel outgrow the instruction cache)
        substep_tile<T, NT, NAPP, GEN, CW, batch_copy_waves<T>(CW, NT), true, PAIR, false>(
            a, u - ent[lo].unit0, smem_raw);
    }
}

__global__ void fill_kernel(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T, int NT, int NAPP, bool GEN, int CW, bool PAIR, bool WOUT = false>
static hipError_t launch_v(const SubstepArgs<T> &a, size_t lds_bytes, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)substep_kernel<T, NT, NAPP, GEN, CW, PAIR, WOUT>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int nb = ((a.n_tiles + 7) / 8) * 8;
    hipLaunchKernelGGL((substep_kernel<T, NT, NAPP, GEN, CW, PAIR, WOUT>), dim3(nb), dim3(kthreads(CW)),
                       lds_bytes, s, a);
    return hipGetLastError();
}

template <typename T, int NT, int CW, bool PAIR>
static hipError_t launch_nt(const SubstepArgs<T> &a, bool gen, size_t lds_bytes, hipStream_t s) {
    if (a.wout) {
        // W+Q output: the variants without global-memory coefficients only
        if (gen) return hipErrorInvalidValue;
        return a.napp == 1 ? launch_v<T, NT, 1, false, CW, PAIR, true>(a, lds_bytes, s)
                           : launch_v<T, NT, 2, false, CW, PAIR, true>(a, lds_bytes, s);
    }
    if (a.napp == 1)
        return gen ? launch_v<T, NT, 1, true, CW, PAIR>(a, lds_bytes, s)
                   : launch_v<T, NT, 1, false, CW, PAIR>(a, lds_bytes, s);
    return gen ? launch_v<T, NT, 2, true, CW, PAIR>(a, lds_bytes, s)
               : launch_v<T, NT, 2, false, CW, PAIR>(a, lds_bytes, s);
}

// Two cells per thread: FP32 state (the only variants built)
template <typename T>
static constexpr bool pair_built(int cw) {
    return sizeof(T) == 4 && (cw == 4 || cw == 8);
}

template <typename T>
hipError_t launch_substep(const SubstepArgs<T> &a, int nt, bool gen, int cw, size_t lds_bytes,
                          hipStream_t s) {
    if (a.ncf > MAX_CFIELDS || (cw != 4 && cw != 8) || a.chunks > CPS_MAX ||
        (a.napp != 1 && a.napp != 2) || (a.pair && !pair_built<T>(cw)))
        return hipErrorInvalidValue;
    // terms per record: 8 (2D P_N: up to four per axis), 10 (3D P7), 12
    if (a.pair) {
        if constexpr (sizeof(T) == 4) {
            if (cw == 8)
                return nt <= 8    ? launch_nt<T, 8, 8, true>(a, gen, lds_bytes, s)
                       : nt <= 10 ? launch_nt<T, 10, 8, true>(a, gen, lds_bytes, s)
                                  : launch_nt<T, MAXT, 8, true>(a, gen, lds_bytes, s);
            return nt <= 8    ? launch_nt<T, 8, 4, true>(a, gen, lds_bytes, s)
                   : nt <= 10 ? launch_nt<T, 10, 4, true>(a, gen, lds_bytes, s)
                              : launch_nt<T, MAXT, 4, true>(a, gen, lds_bytes, s);
        }
        return hipErrorInvalidValue;
    }
    if (cw == 8)
        return nt <= 8    ? launch_nt<T, 8, 8, false>(a, gen, lds_bytes, s)
               : nt <= 10 ? launch_nt<T, 10, 8, false>(a, gen, lds_bytes, s)
                          : launch_nt<T, MAXT, 8, false>(a, gen, lds_bytes, s);
    return nt <= 8    ? launch_nt<T, 8, 4, false>(a, gen, lds_bytes, s)
           : nt <= 10 ? launch_nt<T, 10, 4, false>(a, gen, lds_bytes, s)
                      : launch_nt<T, MAXT, 4, false>(a, gen, lds_bytes, s);
}

// (the dynamic LDS limit of batch_kernel is raised by batch_occupancy, which
// runs before any launch and outside graph capture)
template <typename T, int NAPP, int CW, bool PAIR>
static const void *batch_fn_cw(int nt, bool gen) {
    if (nt <= 8)
        return gen ? (const void *)batch_kernel<T, 8, NAPP, true, CW, PAIR>
                   : (const void *)batch_kernel<T, 8, NAPP, false, CW, PAIR>;
    if (nt <= 10)
        return gen ? (const void *)batch_kernel<T, 10, NAPP, true, CW, PAIR>
                   : (const void *)batch_kernel<T, 10, NAPP, false, CW, PAIR>;
    return gen ? (const void *)batch_kernel<T, MAXT, NAPP, true, CW, PAIR>
               : (const void *)batch_kernel<T, MAXT, NAPP, false, CW, PAIR>;
}

template <typename T>
static const void *batch_fn(int nt, bool gen, int cw, int napp, bool pair) {
    if (pair) {
        if constexpr (sizeof(T) == 4) {
            if (cw == 8)
                return napp == 2 ? batch_fn_cw<T, 2, 8, true>(nt, gen) : batch_fn_cw<T, 1, 8, true>(nt, gen);
            return napp == 2 ? batch_fn_cw<T, 2, 4, true>(nt, gen) : batch_fn_cw<T, 1, 4, true>(nt, gen);
        }
        return nullptr;
    }
    if (napp == 2)
        return cw == 8 ? batch_fn_cw<T, 2, 8, false>(nt, gen) : batch_fn_cw<T, 2, 4, false>(nt, gen);
    return cw == 8 ? batch_fn_cw<T, 1, 8, false>(nt, gen) : batch_fn_cw<T, 1, 4, false>(nt, gen);
}

template <typename T>
hipError_t launch_batch(const BatchLaunch<T> *bl, int nt, bool gen, int cw, int napp, bool pair,
                        int grid, size_t lds_bytes, hipStream_t s) {
    if (grid < 1 || lds_bytes > 160 * 1024 - 64 || (cw != 4 && cw != 8) || (napp != 1 && napp != 2) ||
        (pair && !pair_built<T>(cw)))
        return hipErrorInvalidValue;
    const void *f = batch_fn<T>(nt, gen, cw, napp, pair);
    if (!f) return hipErrorInvalidValue;
    void *args[] = {(void *)&bl};
    return hipLaunchKernel(f, dim3(grid), dim3(batch_threads<T>(cw, nt)), args, lds_bytes, s);
}

template <typename T>
hipError_t batch_occupancy(int nt, bool gen, int cw, int napp, bool pair, size_t lds_bytes,
                           int *blocks) {
    if ((cw != 4 && cw != 8) || (napp != 1 && napp != 2) || (pair && !pair_built<T>(cw)))
        return hipErrorInvalidValue;
    const void *f = batch_fn<T>(nt, gen, cw, napp, pair);
    if (!f) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, f, batch_threads<T>(cw, nt), lds_bytes);
}

hipError_t launch_fill(int *p, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void decay_kernel(const T *sig, T *E, int64_t count, double dt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double x = -(double)sig[i] * dt / 2;
        double y;
        if (fabs(x) > 2e-4)
            y = (exp(x) - 1) / x;
        else
            y = 1 + x * .5 + (x * x) / 6;
        E[i] = (T)y;
    }
}

template <typename T>
hipError_t launch_decay(const T *sig, T *E, int64_t count, double dt, hipStream_t s) {
    int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(decay_kernel<T>, dim3(blocks), dim3(256), 0, s, sig, E, count, dt);
    return hipGetLastError();
}

template <typename T>
__global__ void scale_kernel(T *U, const T *F, T F_s, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        U[i] = U[i] * (F ? F[i] : F_s);
}

template <typename T>
hipError_t launch_scale(T *U, const T *F, double F_s, int64_t count, hipStream_t s) {
    int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, s, U, F, (T)F_s, count);
    return hipGetLastError();
}

template <typename T>
__global__ void interp_kernel(const T *U0, const T *U, double *out, int64_t count,
                              double lambda) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (1 - lambda) * (double)U0[i] + lambda * (double)U[i];
}

template <typename T>
hipError_t launch_interp(const T *U0, const T *U, double *out, int64_t count, double lambda,
                         hipStream_t s) {
    int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(interp_kernel<T>, dim3(blocks), dim3(256), 0, s, U0, U, out, count,
                       lambda);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double trap_w(int i, int ext, int stag) {
    // starmap_2d_ex_l2norm.m:96-97: interior weight 1, end weights s/2 with
    // s = (extent == n); an axis of extent 1 (no extent along it) weighs 1
    if (ext == 1) return 1.0;
    if (i == 0 || i == ext - 1) return stag ? 0.0 : 0.5;
    return 1.0;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(256) void reduce_kernel(const T *U, ReduceGeom g, double *part) {
    __shared__ double red[3][4];
    const int64_t total = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    double s = 0, q = 0, bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int i0 = (int)(i % g.ext[0]);
        const int64_t r = i / g.ext[0];
        const int i1 = (int)(r % g.ext[1]);
        const int i2 = (int)(r / g.ext[1]);
        const double v = (double)U[i0 + (int64_t)g.pitch0 * (i1 + (int64_t)g.ext1 * i2)];
        const double w = trap_w(i0, g.ext[0], g.stag[0]) * trap_w(i1, g.ext[1], g.stag[1]) *
                         trap_w(i2, g.ext[2], g.stag[2]);
        s += v;
        q += w * v * v;
        bad += isfinite(v) ? 0.0 : 1.0;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    bad = wave_sum(bad);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = s;
        red[1][wv] = q;
        red[2][wv] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            a += red[0][k];
            b += red[1][k];
            c += red[2][k];
        }
        part[3 * blockIdx.x + 0] = a;
        part[3 * blockIdx.x + 1] = b * g.cell;
        part[3 * blockIdx.x + 2] = c;
    }
}

template <typename T>
hipError_t launch_reduce(const T *U, const ReduceGeom &g, double *part, int blocks,
                         hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel<T>, dim3(blocks), dim3(256), 0, s, U, g, part);
    return hipGetLastError();
}

__global__ void reduce_final_kernel(const double *part, int blocks, double *out3) {
    // one wavefront, fixed order: deterministic result
    double a = 0, b = 0, c = 0;
    for (int k = threadIdx.x; k < blocks; k += 64) {
        a += part[3 * k];
        b += part[3 * k + 1];
        c += part[3 * k + 2];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (threadIdx.x == 0) {
        out3[0] = a;
        out3[1] = b;
        out3[2] = c;
    }
}

hipError_t launch_reduce_final(const double *part, int blocks, double *out3, hipStream_t s) {
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(64), 0, s, part, blocks, out3);
    return hipGetLastError();
}

#define SMP_INSTANTIATE(T)                                                                     \
    template hipError_t launch_substep<T>(const SubstepArgs<T> &, int, bool, int, size_t,        \
                                          hipStream_t);                                         \
    template hipError_t launch_decay<T>(const T *, T *, int64_t, double, hipStream_t);        \
    template hipError_t launch_scale<T>(T *, const T *, double, int64_t, hipStream_t);        \
    template hipError_t launch_interp<T>(const T *, const T *, double *, int64_t, double,     \
                                         hipStream_t);                                         \
    template hipError_t launch_reduce<T>(const T *, const ReduceGeom &, double *, int,         \
                                         hipStream_t);                                         \
    template hipError_t launch_batch<T>(const BatchLaunch<T> *, int, bool, int, int, bool, int,  \
                                        size_t, hipStream_t);                                  \
    template hipError_t batch_occupancy<T>(int, bool, int, int, bool, size_t, int *);

SMP_INSTANTIATE(double)
SMP_INSTANTIATE(float)

}  // namespace smp
