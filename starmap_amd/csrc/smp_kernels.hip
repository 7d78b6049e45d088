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

// One workgroup = one spatial tile of up to 256 cells, one thread per cell.
// Source tiles go from HBM straight into LDS slots (global_load_lds, no
// registers), the group records the same way.  While the targets of group g
// are updated, the source tiles and target values of group g+1 and the
// records of group g+2 are in flight; one barrier per group.  The schedule
// never lets group g+1 copy into a slot that group g reads.
template <typename T, int NT>
__global__ __launch_bounds__(256) void substep_kernel(const SubstepArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char *ring = smem_raw;
    T *cfa = reinterpret_cast<T *>(smem_raw + 3 * BLOB_SZ);
    T *lds = cfa + a.ncf * 256;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.n_tiles) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    int b = tile;
    const int t0 = b % a.ntile[0];
    b /= a.ntile[0];
    const int t1 = b % a.ntile[1];
    const int t2 = b / a.ntile[1];
    const int o0 = t0 * a.tile[0], o1 = t1 * a.tile[1], o2 = t2 * a.tile[2];
    const int ncell = a.tile[0] * a.tile[1] * a.tile[2];
    const bool has_cell = tid < ncell;
    const int c0 = tid % a.tile[0];
    const int c1 = (tid / a.tile[0]) % a.tile[1];
    const int c2 = tid / (a.tile[0] * a.tile[1]);
    const int g0 = o0 + c0, g1 = o1 + c1, g2 = o2 + c2;
    const bool in_store = has_cell && g0 < a.E[0] && g1 < a.E[1] && g2 < a.E[2];
    const int lbase = c0 + a.lds[0] * (c1 + a.lds[1] * c2);
    const int64_t gidx = g0 + (int64_t)a.pitch0 * (g1 + (int64_t)a.ext1 * g2);
    const int S = a.slot_elems;
    const int L1 = a.lds[1];
    const T half_dt = (T)a.half_dt;
    const int ng = a.ngroups;
    // elements per DMA chunk and dwords per element
    constexpr int DPE = sizeof(T) / 4;
    constexpr int EPC = 64 / DPE;

    for (int c = 0; c < a.ncf; ++c)
        cfa[c * 256 + tid] = in_store ? reinterpret_cast<const T *>(a.cfields[c])[gidx] : (T)0;

    auto hdr_nl = [&](int buf) { return reinterpret_cast<const int *>(ring + buf * BLOB_SZ)[0]; };
    auto hdr_nt = [&](int buf) { return reinterpret_cast<const int *>(ring + buf * BLOB_SZ)[1]; };
    auto recs = [&](int buf) {
        return reinterpret_cast<const TgtRec *>(ring + buf * BLOB_SZ + BLOB_REC_OFF);
    };
    auto ldrec = [&](int buf) {
        return reinterpret_cast<const DevLoad *>(ring + buf * BLOB_SZ + BLOB_LD_OFF);
    };
    // HBM -> LDS: the blob of group gq (three 1 KB chunks, 16 bytes a lane)
    auto blob_dma = [&](int gq, int buf) {
        if (wave < 3) {
            const unsigned char *src = a.blobs + (size_t)gq * BLOB_SZ + wave * 1024 + lane * 16;
            __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(ring + buf * BLOB_SZ + wave * 1024),
                                             16, 0, 0);
        }
    };
    // HBM -> LDS: the source tiles of the group in blob buf, 4 bytes a lane;
    // chunk c of the group is slot-local dword range [64*cc, 64*cc+64) of
    // load li, copied by wave c % 4
    auto src_dma = [&](int buf) {
        const int nl = hdr_nl(buf);
        const DevLoad *L = ldrec(buf);
        const int cps = a.chunks;
        const int total = nl * cps;
        for (int c = wave; c < total; c += 4) {
            const int li = c / cps;
            const int cc = c - li * cps;
            const int e = cc * EPC + lane / DPE;
            if (e < S) {
                const DevLoad &ld = L[li];
                const int l0 = e & (a.lds[0] - 1);
                const int row = e >> a.lds0_shift;
                const int l2 = (int)(((float)row + 0.5f) * a.inv_lds1);
                const int l1 = row - l2 * L1;
                const int x0 = wrap_or_clamp(o0 + l0 + ld.org[0], a.n[0], a.bc[0], ld.ext[0]);
                const int x1 = wrap_or_clamp(o1 + l1 + ld.org[1], a.n[1], a.bc[1], ld.ext[1]);
                const int x2 = wrap_or_clamp(o2 + l2 + ld.org[2], a.n[2], a.bc[2], ld.ext[2]);
                const T *src = a.U + (int64_t)ld.comp * a.slab + x0 +
                               (int64_t)a.pitch0 * (x1 + (int64_t)a.ext1 * x2);
                const int slot = __builtin_amdgcn_readfirstlane(ld.slot);
                unsigned char *dst = reinterpret_cast<unsigned char *>(lds + slot * S) + cc * 256;
                __builtin_amdgcn_global_load_lds(
                    reinterpret_cast<const unsigned *>(src) + (DPE == 2 ? (lane & 1) : 0),
                    (lds_ptr_t)dst, 4, 0, 0);
            }
        }
    };
    // HBM -> registers: the target values of the group in blob buf
    auto load_targets = [&](int buf, T *u) {
        const TgtRec *r = recs(buf);
        const int nt = __builtin_amdgcn_readfirstlane(hdr_nt(buf));
#pragma unroll
        for (int i = 0; i < GMAX; ++i) {
            if (i < nt) {
                const int4 h0 = reinterpret_cast<const int4 *>(r + i)[0];   // comp ..
                const int4 h1 = reinterpret_cast<const int4 *>(r + i)[1];   // Q_off, ext[3]
                if (has_cell && g0 < h1.y && g1 < h1.z && g2 < h1.w)
                    u[i] = a.U[(int64_t)h0.x * a.slab + gidx];
            }
        }
    };
    auto coef = [&](int off, const void *p, double s) -> T {
        if (off >= 0) return cfa[off + tid];
        if (off == -1) return (T)s;
        return in_store ? reinterpret_cast<const T *>(p)[gidx] : (T)0;
    };

    T u_cur[GMAX], u_nxt[GMAX];
    if (ng > 0) blob_dma(0, 0);
    if (ng > 1) blob_dma(1, 1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (ng > 0) {
        src_dma(0);
        load_targets(0, u_cur);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    for (int gq = 0; gq < ng; ++gq) {
        const int buf = gq % 3;
        if (gq + 1 < ng) {
            src_dma((gq + 1) % 3);
            load_targets((gq + 1) % 3, u_nxt);
        }
        if (gq + 2 < ng) blob_dma(gq + 2, (gq + 2) % 3);
        const TgtRec *rb = recs(buf);
        const int nt = __builtin_amdgcn_readfirstlane(hdr_nt(buf));
#pragma unroll
        for (int i = 0; i < GMAX; ++i) {
            if (i < nt) {
                // one batch of wide LDS reads for the whole record
                const TgtRec r = rb[i];
                const bool act = has_cell && g0 < r.ext[0] && g1 < r.ext[1] && g2 < r.ext[2];
                // W = -sumcell(...): the coefficients carry -M(j,k)/h;
                // padding terms add coef 0 times a zero difference
                T w = 0;
                if (__builtin_amdgcn_readfirstlane(r.nterm) > 0) {
#pragma unroll
                    for (int q = 0; q < NT; ++q) {
                        const int o = r.t[q].off + lbase;
                        w += (T)r.t[q].coef * (lds[o + r.t[q].stride] - lds[o]);
                    }
                }
                const T sig = coef(r.sig_off, r.sig_p, r.sig_s);
                const T E = coef(r.E_off, r.E_p, r.E_s);
                const T wq = r.Q_off == -1 ? w : w + coef(r.Q_off, r.Q_p, 0.0);
                T u = u_cur[i];
                for (int app = 0; app < a.napp; ++app) {
                    // U + dt/2*(W+Q - s.*U).*E   (starmap_solver.m:352-354)
                    u = u + (half_dt * (wq - sig * u)) * E;
                    if (r.filt) {
                        const bool s1 = a.fslot[app] != 0;
                        const void *fp = s1 ? r.F[1] : r.F[0];
                        const double fs = s1 ? r.F_s[1] : r.F_s[0];
                        u = u * (fp && act ? reinterpret_cast<const T *>(fp)[gidx] : (T)fs);
                    }
                }
                if (act) a.U[(int64_t)r.comp * a.slab + gidx] = u;
            }
        }
        // the copies of group g+1 (and its target values) must have landed
        // and every wave must be done with group g before anyone goes on
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < GMAX; ++i) u_cur[i] = u_nxt[i];
    }
}

template <typename T, int NT>
static hipError_t launch_nt(const SubstepArgs<T> &a, int threads, size_t lds_bytes,
                            hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)substep_kernel<T, NT>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int nb = ((a.n_tiles + 7) / 8) * 8;
    hipLaunchKernelGGL((substep_kernel<T, NT>), dim3(nb), dim3(threads), lds_bytes, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_substep(const SubstepArgs<T> &a, int nt, int threads, size_t lds_bytes,
                          hipStream_t s) {
    if (a.ncf > MAX_CFIELDS || threads != 256) return hipErrorInvalidValue;
    if (nt <= 4) return launch_nt<T, 4>(a, threads, lds_bytes, s);
    if (nt <= 8) return launch_nt<T, 8>(a, threads, lds_bytes, s);
    return launch_nt<T, MAXT>(a, threads, lds_bytes, s);
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
    template hipError_t launch_substep<T>(const SubstepArgs<T> &, int, int, size_t, hipStream_t); \
    template hipError_t launch_decay<T>(const T *, T *, int64_t, double, hipStream_t);        \
    template hipError_t launch_scale<T>(T *, const T *, double, int64_t, hipStream_t);        \
    template hipError_t launch_interp<T>(const T *, const T *, double *, int64_t, double,     \
                                         hipStream_t);                                         \
    template hipError_t launch_reduce<T>(const T *, const ReduceGeom &, double *, int,         \
                                         hipStream_t);

SMP_INSTANTIATE(double)
SMP_INSTANTIATE(float)

}  // namespace smp
