// gfx950 LDS-staged update of the discrete adjoint: the adjoint of one
// launch of the substep kernel for the targets of one parity, with the
// transposed flux of the launch after it fused in.
//
// Per target j and point (starmap_solver.m:352-358 and :404-410 reversed):
//     lambda' = (T)(lambda_j + sum_i c_ij (D^T mu_i))   (the launch after)
//     W       = sum_k c_jk (U_k(hi) - U_k(lo))          (the forward flux)
//     per half-step application, last first, with l = F * lambda':
//         mu_j += a E l,  g += l (-a u E + a (W + Q - s u) dE/ds),
//         lambda' = (1 - a s E) l
// and g summed over the targets of each grid type into dJ/dsigma_a (all
// targets) and dJ/dsigma_s0 (j > 0).  The arithmetic of each target and
// point is that of adj_update_kernel (smp_adjoint_kernels.hip), so both
// kernels give the same lambda (and the same gradients up to the order of
// the sum over targets); this one reads every source tile from HBM once per
// workgroup tile instead of gathering each sample from global memory.
//
// Both fluxes are sums of coef * (hi - lo) over LDS tiles, as in the
// substep kernel: W over the state tiles of the sources, the transposed
// flux over their mu tiles with the forward geometry and coefficient -c_ij.
// D^T mu is the forward-form difference of mu extended by zero.  A
// cell-centred mu tile reads index n (the unused layer of a cell-centred
// slab, which stays zero) where the forward would clamp at the low or the
// high end.  The boundary faces of an extrapolation axis carry no flux along
// it (the forward difference there is identically zero): a face-centred mu
// tile used for a coupling along such an axis e is staged with its boundary
// faces of e read from a zero buffer (the same mu still counts along the
// other axes, through a tile staged without that mask).
//
// Workgroup: 4*ADJ_PARTS cell waves in ADJ_PARTS parts of 256 threads (one
// cell each; each part updates four of a group's targets, consecutive in the
// order of grid types) beside the copy waves that make up 16 waves, one
// workgroup per CU.  Each part sums a point's gradient over its targets of a
// grid type in order; the parts' sums are added in a fixed order at the end
// of the tile.  The tile shape is
// compiled in per number of kernel axes (adj_tile_dim), so the high sample
// of a term is an immediate offset from the low one.
#include "smp_kernels.h"

#include <cmath>

namespace smp {

namespace {

#define SMP_GLOBAL __attribute__((address_space(1)))
#define SMP_CONST __attribute__((address_space(4)))
template <typename P>
__device__ __forceinline__ SMP_GLOBAL P *glb(P *p) {
    return (SMP_GLOBAL P *)p;
}

typedef __attribute__((address_space(3))) void *lds_ptr_t;

__device__ __forceinline__ int wrap_or_clamp(int i, int n, int bc, int ext) {
    if (bc == 0) {
        if (i < 0) i += n;
        if (i >= n) i %= n;
        return i;
    }
    return i < 0 ? 0 : (i >= ext ? ext - 1 : i);
}

// workgroups are dealt round-robin to the 8 XCDs: a contiguous run of tiles
// per XCD, so that halo neighbours share an L2
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

constexpr int ACW = 4 * ADJ_PARTS;     // cell waves
constexpr int ANCW = 16 - ACW;         // copy waves
constexpr int ANCT = ACW * 64;         // cell threads
constexpr int AHALF = 256;             // cells per tile, cell threads per part
constexpr int AP = ADJ_PARTS;
static_assert(ADJ_THREADS == (ACW + ANCW) * 64, "adjoint workgroup");
static_assert(ADJ_BLOB_LD_OFF + ADJ_MAXL * (int)sizeof(DevLoad) <= BLOB_SZ, "adjoint blob layout");
static_assert(sizeof(AdjRec) % 16 == 0, "adjoint record layout");

}  // namespace

size_t adj_lds_overhead(int ncf, int ngt, int elem) {
    return 3 * (size_t)BLOB_SZ + (size_t)ncf * AHALF * elem + 2 * AP * (size_t)ngt * AHALF * sizeof(double);
}

template <typename T, int NDIM, bool FUSE, bool GEN, bool HASW>
__global__ __launch_bounds__(ADJ_THREADS) void adj_lds_kernel(const AdjTileArgs<T> a) {
    constexpr int NT = ADJ_TERMS;
    constexpr int TL0 = adj_tile_dim(NDIM, 0) + 1;   // LDS slot extents
    constexpr int TL1 = NDIM >= 2 ? adj_tile_dim(NDIM, 1) + 1 : 1;
    constexpr int TX0 = adj_tile_dim(NDIM, 0), TX1 = adj_tile_dim(NDIM, 1), TX2 = adj_tile_dim(NDIM, 2);
    // byte distance of the high sample from the low one, per kernel axis
    constexpr int SB[3] = {(int)sizeof(T), TL0 * (int)sizeof(T), TL0 * TL1 * (int)sizeof(T)};
    static_assert(TL0 * TX1 * TX2 <= AHALF, "one cell per thread of a half");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.n_tiles) return;
    unsigned char *ring = smem_raw;
    T *cfa = reinterpret_cast<T *>(smem_raw + 3 * BLOB_SZ);
    double *gacc = reinterpret_cast<double *>(cfa + a.ncf * AHALF);
    T *lds = reinterpret_cast<T *>(gacc + 2 * AP * a.ngt * AHALF);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool copier = __builtin_amdgcn_readfirstlane(tid) >= ANCT;
    const int ctid = tid & (AHALF - 1);
    const int half = __builtin_amdgcn_readfirstlane(min(tid >> 8, AP - 1));   // the part of a cell thread
    const double hd = a.half_dt;
    int b = tile;
    const int t0 = b % a.ntile[0];
    b /= a.ntile[0];
    const int t1 = b % a.ntile[1];
    const int t2 = b / a.ntile[1];
    const int o0 = t0 * TX0, o1 = t1 * TX1, o2 = t2 * TX2;
    // cell threads by LDS slot column: a wave's LDS reads are consecutive
    const int rr = ctid / TL0;
    const int c0 = ctid % TL0, c1 = rr % TX1, c2 = rr / TX1;
    const bool has_cell = !copier && c0 < TX0 && c2 < TX2;
    const int g0 = o0 + c0, g1 = o1 + c1, g2 = o2 + c2;
    const bool in_store = has_cell && g0 < a.E[0] && g1 < a.E[1] && g2 < a.E[2];
    const int lbase = min(c0, TX0 - 1) + TL0 * (min(c1, TX1 - 1) + TL1 * min(c2, TX2 - 1));
    const int64_t gidx = g0 + (int64_t)a.pitch0 * (g1 + (int64_t)a.ext1 * g2);
    constexpr int TL2 = NDIM >= 3 ? adj_tile_dim(NDIM, 2) + 1 : 1;
    constexpr int S = TL0 * TL1 * TL2;        // elements per LDS slot
    const int ng = a.ngroups;
    constexpr int DPE = sizeof(T) / 4;
    constexpr int EPC = 64 / DPE;
    constexpr int CH = (S * DPE + 63) / 64;   // 64-lane copies per slot
    static_assert(CH <= CPS_MAX, "copies per slot");

    auto hdr_nl = [&](int buf) { return reinterpret_cast<const int *>(ring + buf * BLOB_SZ)[0]; };
    auto ldrec = [&](int buf) {
        return reinterpret_cast<const DevLoad *>(ring + buf * BLOB_SZ + ADJ_BLOB_LD_OFF);
    };

    // ---- copy waves: per-tile offsets of the element each lane copies ----
    // for a face-centred tile (wrap / clamp resolved); per chunk and axis a
    // bit for lanes at coordinate 0 (zb), past the last face (hb) and on a
    // boundary face (fb) of an extrapolation axis.  A cell-centred state
    // tile looks one cell lower, x-1, clamped at 0 (periodic: n-1); a
    // cell-centred mu tile reads index n (zero) below 0 and past the last
    // cell; a masked mu tile reads the zero buffer on the boundary faces of
    // its axis.
    int po[CH];
    int zb[3] = {0, 0, 0}, hb[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
    const int st[3] = {DPE, a.pitch0 * DPE, a.pitch0 * a.ext1 * DPE};
    int wd[3], wdm[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        wd[e] = a.bc[e] == 0 ? (a.n[e] - 1) * st[e] : 0;
        wdm[e] = a.bc[e] == 0 ? (a.n[e] - 1) * st[e] : a.n[e] * st[e];
    }
    bool low_inner = false, mu_inner = false;
    if (copier) {
        const int oo[3] = {o0, o1, o2};
#pragma unroll
        for (int cc = 0; cc < CH; ++cc) {
            const int e = cc * EPC + lane / DPE;
            po[cc] = -1;
            if (e < S) {
                int l[3];
                l[0] = e % TL0;
                l[1] = (e / TL0) % TL1;
                l[2] = e / (TL0 * TL1);
                int x[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    x[q] = wrap_or_clamp(oo[q] + l[q], a.n[q], a.bc[q], a.n[q] + a.bc[q]);
                    zb[q] |= (x[q] == 0) << cc;
                    hb[q] |= (a.bc[q] == 1 && oo[q] + l[q] > a.n[q]) << cc;
                    fb[q] |= (a.bc[q] == 1 && (x[q] == 0 || x[q] == a.n[q])) << cc;
                }
                po[cc] = (x[0] + a.pitch0 * (x[1] + a.ext1 * x[2])) * DPE + (DPE == 2 ? (lane & 1) : 0);
            }
        }
        int zany = 0, hany = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e)
            if (a.lds[e] > 1) {
                zany |= zb[e];
                hany |= hb[e];
            }
        low_inner = !__any(zany != 0);
        mu_inner = !__any((zany | hany) != 0);
    }
    auto blob_dma = [&](int gq, int buf) {
        if (tid >= ANCT + 64) return;
#pragma unroll
        for (int k = 0; k < BLOB_SZ / 1024; ++k) {
            SMP_GLOBAL const unsigned char *src = glb(a.blobs) + (size_t)gq * BLOB_SZ + k * 1024 + lane * 16;
            __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(ring + buf * BLOB_SZ + k * 1024), 16, 0, 0);
        }
    };
    SMP_GLOBAL const unsigned *zl = glb(a.zeros) + lane;
    auto src_dma = [&](int buf) {
        // comp | slot << 20 | (cell-centred axes) << 28 | mode << 29...: the
        // mode (0 state, 1 mu, 2+e mu masked on e) in bits 29-31
        const int nl = __builtin_amdgcn_readfirstlane(hdr_nl(buf));
        const int cw = (tid >> 6) - ACW;
        unsigned myrec = 0;
        if (lane < nl) {
            const DevLoad &d = ldrec(buf)[lane];
            myrec = (unsigned)d.comp | ((unsigned)d.slot << 18) |
                    ((unsigned)((d.org[0] & 1) | ((d.org[1] & 1) << 1) | ((d.org[2] & 1) << 2)) << 26) |
                    ((unsigned)(d.pad[0] & 7) << 29);
        }
        for (int li = cw; li < nl; li += ANCW) {
            const unsigned rec = (unsigned)__builtin_amdgcn_readlane((int)myrec, li);
            const int comp = rec & 0x3FFFF;
            const int slot = (rec >> 18) & 0xFF;
            const int m = (rec >> 26) & 7;
            const int mode = (int)(rec >> 29);
            // without the transposed flux the mu tiles are not read
            if (mode != 0 && !FUSE) continue;
            SMP_GLOBAL const unsigned *cb = glb(reinterpret_cast<const unsigned *>(
                (mode != 0 ? (const T *)a.mu : a.U) + (int64_t)comp * a.slab));
            unsigned char *dst = reinterpret_cast<unsigned char *>(lds + slot * S);
            if (mode <= 1 && (mode ? mu_inner : low_inner)) {
                SMP_GLOBAL const unsigned *cbs = cb - ((m & 1) ? st[0] : 0) - ((m & 2) ? st[1] : 0) -
                                                 ((m & 4) ? st[2] : 0);
#pragma unroll
                for (int cc = 0; cc < CH; ++cc)
                    if (po[cc] >= 0)
                        __builtin_amdgcn_global_load_lds(cbs + po[cc], (lds_ptr_t)(dst + cc * 256), 4, 0, 0);
            } else if (mode == 0) {
#pragma unroll
                for (int cc = 0; cc < CH; ++cc) {
                    if (po[cc] >= 0) {
                        int off = po[cc];
#pragma unroll
                        for (int e = 0; e < 3; ++e)
                            if (m & (1 << e)) off += ((zb[e] >> cc) & 1) ? wd[e] : -st[e];
                        __builtin_amdgcn_global_load_lds(cb + off, (lds_ptr_t)(dst + cc * 256), 4, 0, 0);
                    }
                }
            } else {
                // mask axis of a masked tile (a bit per axis; 0: none)
                const int mb = mode >= 2 ? 1 << (mode - 2) : 0;
#pragma unroll
                for (int cc = 0; cc < CH; ++cc) {
                    if (po[cc] >= 0) {
                        int off = po[cc];
                        bool zero = false;
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            if (m & (1 << e))
                                off += ((zb[e] >> cc) & 1) ? wdm[e] : (((hb[e] >> cc) & 1) ? 0 : -st[e]);
                            if (mb & (1 << e)) zero |= ((fb[e] >> cc) & 1) != 0;
                        }
                        __builtin_amdgcn_global_load_lds(zero ? zl : cb + off, (lds_ptr_t)(dst + cc * 256),
                                                         4, 0, 0);
                    }
                }
            }
        }
    };

    if (copier) {
        if (ng > 0) blob_dma(0, 0);
        if (ng > 1) blob_dma(1, 1);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (ng > 0) src_dma(0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int gq = 0; gq < ng; ++gq) {
            if (gq + 1 < ng) src_dma((gq + 1) % 3);
            if (gq + 2 < ng) blob_dma(gq + 2, (gq + 2) % 3);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        __syncthreads();
        return;
    }

    // ---- cell waves --------------------------------------------------------
    // every lane issues every target load and store (lanes without a target
    // point read element 0 and write to the sink), so the compiler waits for
    // a prefetched value with an exact count
    const int64_t pix = (in_store ? gidx : 0) * (int64_t)sizeof(T);
    SMP_GLOBAL const char *u0_ld = glb(reinterpret_cast<const char *>(a.U0)) + pix;
    SMP_GLOBAL const char *lam_ld = glb(reinterpret_cast<const char *>(a.lam)) + pix;
    SMP_GLOBAL char *lam_st = glb(reinterpret_cast<char *>(a.lam)) + gidx * sizeof(T);
    SMP_GLOBAL char *mu_st = glb(reinterpret_cast<char *>(a.mu)) + gidx * sizeof(T);
    SMP_GLOBAL T *sink = glb(a.sink);
    const int fs0 = __builtin_amdgcn_readfirstlane(a.fslot[0]) ? 1 : 0;
    const int fs1 = __builtin_amdgcn_readfirstlane(a.fslot[1]) ? 1 : 0;
    const int napp = __builtin_amdgcn_readfirstlane(a.napp);
    auto rec_of = [&](int gq) {
        return (const SMP_CONST AdjRec *)(a.recs) + (int64_t)gq * ADJ_RECS + half * ADJ_TPH;
    };
    SMP_GLOBAL const char *w_ld = HASW ? glb(reinterpret_cast<const char *>(a.W)) + pix : nullptr;
    auto load_targets = [&](int gq, T *u0, T *lm, T *wv) {
        const auto r = rec_of(gq);
#pragma unroll
        for (int i = 0; i < ADJ_TPH; ++i) {
            u0[i] = *reinterpret_cast<SMP_GLOBAL const T *>(u0_ld + r[i].uoff);
            lm[i] = *reinterpret_cast<SMP_GLOBAL const T *>(lam_ld + r[i].uoff);
            if (HASW) wv[i] = *reinterpret_cast<SMP_GLOBAL const T *>(w_ld + r[i].uoff);
        }
    };
    const int lb = lbase * (int)sizeof(T);
    // sum over a list of terms by kernel axis: per axis a partial sum of its
    // terms, then the axes in order
    auto flux = [&](const SMP_CONST int32_t *off, const SMP_CONST double *coef) -> double {
        double w = 0;
#pragma unroll
        for (int e = 0; e < NDIM; ++e) {
            double p = 0;
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const unsigned char *lo = smem_raw + (off[NT * e + q] + lb);
                // the difference in the state precision (exact for a zero
                // sample), the sum in FP64
                const T d = *reinterpret_cast<const T *>(lo + SB[e]) - *reinterpret_cast<const T *>(lo);
                p += coef[NT * e + q] * (double)d;
            }
            w += p;
        }
        return w;
    };
    // a coefficient: staged in LDS (off >= 0), a scalar (-1) or, in the GEN
    // variant only, a field of the global table (-2-k; more fields than the
    // LDS stages).  The staged value is read whatever the kind, so that no
    // branch waits for the LDS reads in flight.
    auto coef = [&](int off, double sc) -> double {
        const double staged = (double)cfa[(off > 0 ? off : 0) + ctid];
        double v = off >= 0 ? staged : sc;
        if (GEN && off < -1) {
            const T *p = reinterpret_cast<const T *>(a.gfields[-2 - off]);
            v = (double)glb(p)[in_store ? gidx : 0];
        }
        return v;
    };
    double ga = 0, gs = 0;
    auto update = [&](int gq, const T *u0, const T *lm, const T *wv) {
        const auto rb = rec_of(gq);
#pragma unroll
        for (int i = 0; i < ADJ_TPH; ++i) {
            const auto &r = rb[i];
            const bool act = has_cell & (g0 < r.ext[0]) & (g1 < r.ext[1]) & (g2 < r.ext[2]);
            const double w = HASW ? 0.0 : flux(r.off, r.coef);
            double lam = (double)lm[i];
            if (FUSE) lam = (double)(T)(lam + flux(r.off + MAXT, r.coef + MAXT));
            const double sg = coef(r.sig_off, r.sig_s);
            const double E = coef(r.E_off, r.E_s);
            const double wq = HASW ? (double)wv[i] : w + coef(r.Q_off, 0.0);
            const double dE = coef(r.dE_off, r.dE_s);
            double u[2];
            u[0] = (double)u0[i];
            u[1] = (double)(T)((u[0] + (hd * (wq - sg * u[0])) * E) * r.F_s[fs0]);
            double mu = 0, g = 0;
            if (napp == 2) {
                const double l = r.F_s[fs1] * lam;
                mu += hd * E * l;
                g += l * (-hd * u[1] * E + hd * (wq - sg * u[1]) * dE);
                lam = (1 - hd * sg * E) * l;
            }
            {
                const double l = r.F_s[fs0] * lam;
                mu += hd * E * l;
                g += l * (-hd * u[0] * E + hd * (wq - sg * u[0]) * dE);
                lam = (1 - hd * sg * E) * l;
            }
            SMP_GLOBAL T *dl = act ? reinterpret_cast<SMP_GLOBAL T *>(lam_st + r.uoff) : sink + ctid;
            SMP_GLOBAL T *dm = act ? reinterpret_cast<SMP_GLOBAL T *>(mu_st + r.uoff) : sink + AHALF + ctid;
            *dl = (T)lam;
            *dm = (T)mu;
            ga += g;
            if (r.jpos) gs += g;
            if (r.flush) {
                gacc[(2 * AP * r.gq + 2 * half) * AHALF + ctid] = ga;
                gacc[(2 * AP * r.gq + 2 * half + 1) * AHALF + ctid] = gs;
                ga = 0;
                gs = 0;
            }
        }
    };

    if (half == 0) {
        // coefficient fields of the tile: every load in flight before the
        // first LDS write
        T cv[MAX_CFIELDS];
#pragma unroll
        for (int c = 0; c < MAX_CFIELDS; ++c)
            if (c < a.ncf) cv[c] = glb(reinterpret_cast<const T *>(a.cfields[c]))[in_store ? gidx : 0];
#pragma unroll
        for (int c = 0; c < MAX_CFIELDS; ++c)
            if (c < a.ncf) cfa[c * AHALF + ctid] = in_store ? cv[c] : (T)0;
    }
    for (int e = tid; e < S; e += ANCT) lds[e] = (T)0;   // slot 0: the zero tile
    // gradient sums of the grid types a part has no target of stay zero
    for (int q = 0; q < a.ngt; ++q) {
        gacc[(2 * AP * q + 2 * half) * AHALF + ctid] = 0.0;
        gacc[(2 * AP * q + 2 * half + 1) * AHALF + ctid] = 0.0;
    }
    __syncthreads();
    T u0A[ADJ_TPH], lmA[ADJ_TPH], wA[ADJ_TPH], u0B[ADJ_TPH], lmB[ADJ_TPH], wB[ADJ_TPH];
    if (ng > 0) load_targets(0, u0A, lmA, wA);
    __syncthreads();
    for (int gq = 0; gq < ng; gq += 2) {
        update(gq, u0A, lmA, wA);
        if (gq + 1 < ng) load_targets(gq + 1, u0B, lmB, wB);
        __syncthreads();
        if (gq + 1 >= ng) break;
        update(gq + 1, u0B, lmB, wB);
        if (gq + 2 < ng) load_targets(gq + 2, u0A, lmA, wA);
        __syncthreads();
    }
    __syncthreads();
    // the gradient of each grid type at this point: the sums of the parts
    // in order, added once (part h adds grid types h, h+AP, ...)
    for (int q = half; q < a.ngt; q += AP) {
        if (in_store && g0 < a.gext[q][0] && g1 < a.gext[q][1] && g2 < a.gext[q][2]) {
            double sa = 0, ss = 0;
#pragma unroll
            for (int h = 0; h < AP; ++h) {
                sa += gacc[(2 * AP * q + 2 * h) * AHALF + ctid];
                ss += gacc[(2 * AP * q + 2 * h + 1) * AHALF + ctid];
            }
            const int64_t gp = (int64_t)a.gcode[q] * a.slab + gidx;
            glb(a.Ga)[gp] += sa;
            glb(a.Gs)[gp] += ss;
        }
    }
}

template <typename T, int NDIM, bool FUSE, bool GEN, bool HASW>
static hipError_t launch_adj_v(const AdjTileArgs<T> &a, size_t lds_bytes, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)adj_lds_kernel<T, NDIM, FUSE, GEN, HASW>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int nb = ((a.n_tiles + 7) / 8) * 8;
    hipLaunchKernelGGL((adj_lds_kernel<T, NDIM, FUSE, GEN, HASW>), dim3(nb), dim3(ADJ_THREADS), lds_bytes, s,
                       a);
    return hipGetLastError();
}

template <typename T, int NDIM, bool HASW>
static hipError_t launch_adj_w(const AdjTileArgs<T> &a, bool fuse, bool gen, size_t lds_bytes,
                               hipStream_t s) {
    if (fuse)
        return gen ? launch_adj_v<T, NDIM, true, true, HASW>(a, lds_bytes, s)
                   : launch_adj_v<T, NDIM, true, false, HASW>(a, lds_bytes, s);
    return gen ? launch_adj_v<T, NDIM, false, true, HASW>(a, lds_bytes, s)
               : launch_adj_v<T, NDIM, false, false, HASW>(a, lds_bytes, s);
}

template <typename T, int NDIM>
static hipError_t launch_adj_d(const AdjTileArgs<T> &a, bool fuse, bool gen, size_t lds_bytes,
                               hipStream_t s) {
    return a.W ? launch_adj_w<T, NDIM, true>(a, fuse, gen, lds_bytes, s)
               : launch_adj_w<T, NDIM, false>(a, fuse, gen, lds_bytes, s);
}

template <typename T>
hipError_t launch_adj_lds(const AdjTileArgs<T> &a, int ndim, bool fuse, bool gen, size_t lds_bytes,
                          hipStream_t s) {
    if (a.ncf > MAX_CFIELDS || a.chunks > CPS_MAX || (a.napp != 1 && a.napp != 2) || a.ngt < 1 ||
        a.ngt > 4 || lds_bytes > 160 * 1024 || ndim < 1 || ndim > 3)
        return hipErrorInvalidValue;
    for (int e = 0; e < 3; ++e)
        if (a.tile[e] != adj_tile_dim(ndim, e) || a.lds[e] != (e < ndim ? a.tile[e] + 1 : 1))
            return hipErrorInvalidValue;
    if (a.slot_elems != a.lds[0] * a.lds[1] * a.lds[2]) return hipErrorInvalidValue;
    switch (ndim) {
    case 1: return launch_adj_d<T, 1>(a, fuse, gen, lds_bytes, s);
    case 2: return launch_adj_d<T, 2>(a, fuse, gen, lds_bytes, s);
    default: return launch_adj_d<T, 3>(a, fuse, gen, lds_bytes, s);
    }
}

template hipError_t launch_adj_lds<double>(const AdjTileArgs<double> &, int, bool, bool, size_t,
                                          hipStream_t);
template hipError_t launch_adj_lds<float>(const AdjTileArgs<float> &, int, bool, bool, size_t,
                                         hipStream_t);

}  // namespace smp
