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

// LDS layout: group headers | 2 x (GMAX target records, MAXL load records) |
// source slots
constexpr int BLOB_REC = GMAX * (int)sizeof(TgtRec);
constexpr int BLOB = BLOB_REC + MAXL * (int)sizeof(DevLoad);

size_t substep_lds_overhead() { return MAX_GROUPS * sizeof(DevGroup) + 2 * BLOB; }

// One workgroup = one spatial tile, one thread = one cell.  While the targets
// of group g are updated from LDS and registers, the source tiles and target
// values of group g+1 and the records of group g+2 are in flight; the update
// itself waits on no global load, so the in-order load counter never makes
// it wait for the prefetch.  The coefficient fields the targets use (decay
// rates, decay factors, sources) are staged per tile once.
template <typename T, int NT>
__global__ __launch_bounds__(256) void substep_kernel(const SubstepArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    DevGroup *grp = reinterpret_cast<DevGroup *>(smem_raw);
    unsigned char *blob = smem_raw + MAX_GROUPS * sizeof(DevGroup);
    T *cfa = reinterpret_cast<T *>(blob + 2 * BLOB);
    const int nthr = blockDim.x;
    T *lds = cfa + a.ncf * nthr;
    constexpr int RECV = sizeof(TgtRec) / 16;
    constexpr int LDV = sizeof(DevLoad) / 16;   // DevLoad is 32 bytes
    static_assert(sizeof(DevLoad) == 32, "DevLoad layout");
    static_assert(GMAX * RECV + MAXL * LDV <= 256, "one blob chunk per thread");
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.n_tiles) return;
    const int tid = threadIdx.x;
    int b = tile;
    const int t0 = b % a.ntile[0];
    b /= a.ntile[0];
    const int t1 = b % a.ntile[1];
    const int t2 = b / a.ntile[1];
    const int o0 = t0 * a.tile[0], o1 = t1 * a.tile[1], o2 = t2 * a.tile[2];
    const int c0 = tid % a.tile[0];
    const int c1 = (tid / a.tile[0]) % a.tile[1];
    const int c2 = tid / (a.tile[0] * a.tile[1]);
    const int g0 = o0 + c0, g1 = o1 + c1, g2 = o2 + c2;
    const bool in_store = g0 < a.E[0] && g1 < a.E[1] && g2 < a.E[2];
    const int lbase = c0 + a.lds[0] * (c1 + a.lds[1] * c2);
    const int64_t gidx = g0 + (int64_t)a.pitch0 * (g1 + (int64_t)a.ext1 * g2);
    const int S = a.slot_elems;
    const int L0 = a.lds[0], L1 = a.lds[1];
    const T half_dt = (T)a.half_dt;
    const int ng = a.ngroups;

    for (int i = tid; i < ng; i += nthr) grp[i] = a.groups[i];
    for (int c = 0; c < a.ncf; ++c)
        cfa[c * nthr + tid] = in_store ? reinterpret_cast<const T *>(a.cfields[c])[gidx] : (T)0;

    // the staged elements of this thread: element e = tid + k*nthr of a
    // group's load list is slot-local element r of load li; its tile-local
    // coordinates are fixed for the whole kernel
    int epack[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int e = tid + k * nthr;
        const int li = e / S;
        const int r = e - li * S;
        const int l0 = r % L0;
        const int rr = r / L0;
        // l0 < 1024, l1 < 128, l2 < 128, li < 128 (checked by the host)
        epack[k] = l0 | ((rr % L1) << 10) | ((rr / L1) << 17) | (li << 24);
    }

    T pf[PF];
    int pdst[PF];
    int4 bpf = make_int4(0, 0, 0, 0);
    T u_cur[GMAX], u_nxt[GMAX];

    auto recs = [&](int buf) { return reinterpret_cast<TgtRec *>(blob + buf * BLOB); };
    auto ldr = [&](int buf) { return reinterpret_cast<DevLoad *>(blob + buf * BLOB + BLOB_REC); };

    // global -> registers: the source tiles of a group (load records in LDS)
    auto fetch_src = [&](const DevGroup &gr, int buf) {
        const int nl = gr.l1 - gr.l0;
        const DevLoad *L = ldr(buf);
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int li = epack[k] >> 24;
            pdst[k] = -1;
            if (li < nl) {
                const DevLoad &ld = L[li];
                const int l0 = epack[k] & 1023, l1 = (epack[k] >> 10) & 127,
                          l2 = (epack[k] >> 17) & 127;
                const int x0 = wrap_or_clamp(o0 + l0 + ld.org[0], a.n[0], a.bc[0], ld.ext[0]);
                const int x1 = wrap_or_clamp(o1 + l1 + ld.org[1], a.n[1], a.bc[1], ld.ext[1]);
                const int x2 = wrap_or_clamp(o2 + l2 + ld.org[2], a.n[2], a.bc[2], ld.ext[2]);
                pf[k] = a.U[(int64_t)ld.comp * a.slab + x0 +
                            (int64_t)a.pitch0 * (x1 + (int64_t)a.ext1 * x2)];
                pdst[k] = ld.slot * S + l0 + L0 * (l1 + L1 * l2);
            }
        }
    };
    auto commit_src = [&]() {
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (pdst[k] >= 0) lds[pdst[k]] = pf[k];
    };
    // global -> registers -> LDS: target and load records of a group
    auto fetch_blob = [&](const DevGroup &gr) {
        const int nr = (gr.t1 - gr.t0) * RECV;
        const int nl = (gr.l1 - gr.l0) * LDV;
        if (tid < nr)
            bpf = reinterpret_cast<const int4 *>(a.targets + gr.t0)[tid];
        else if (tid >= GMAX * RECV && tid < GMAX * RECV + nl)
            bpf = reinterpret_cast<const int4 *>(a.loads + gr.l0)[tid - GMAX * RECV];
    };
    auto commit_blob = [&](const DevGroup &gr, int buf) {
        const int nr = (gr.t1 - gr.t0) * RECV;
        const int nl = (gr.l1 - gr.l0) * LDV;
        if (tid < nr || (tid >= GMAX * RECV && tid < GMAX * RECV + nl))
            reinterpret_cast<int4 *>(blob + buf * BLOB)[tid] = bpf;
    };
    // global -> registers: the target values of a group
    auto load_targets = [&](const DevGroup &gr, int buf, T *u) {
        const TgtRec *r = recs(buf);
        const int nt = gr.t1 - gr.t0;
#pragma unroll
        for (int i = 0; i < GMAX; ++i) {
            if (i < nt) {
                const TgtRec &ri = r[i];
                if (g0 < ri.ext[0] && g1 < ri.ext[1] && g2 < ri.ext[2])
                    u[i] = a.U[(int64_t)ri.comp * a.slab + gidx];
            }
        }
    };
    auto coef = [&](int off, const void *p, double s) -> T {
        if (off >= 0) return cfa[off + tid];
        if (off == -1) return (T)s;
        return reinterpret_cast<const T *>(p)[gidx];
    };

    __syncthreads();
    if (ng > 0) {
        fetch_blob(grp[0]);
        commit_blob(grp[0], 0);
        if (ng > 1) {
            fetch_blob(grp[1]);
            commit_blob(grp[1], 1);
        }
    }
    __syncthreads();
    if (ng > 0) {
        fetch_src(grp[0], 0);
        commit_src();
        load_targets(grp[0], 0, u_cur);
    }
    __syncthreads();

    for (int gq = 0; gq < ng; ++gq) {
        const DevGroup gr = grp[gq];
        const int buf = gq & 1;
        if (gq + 1 < ng) {
            load_targets(grp[gq + 1], buf ^ 1, u_nxt);
            fetch_src(grp[gq + 1], buf ^ 1);
        }
        if (gq + 2 < ng) fetch_blob(grp[gq + 2]);
        const TgtRec *rb = recs(buf);
        const int nt = gr.t1 - gr.t0;
#pragma unroll
        for (int i = 0; i < GMAX; ++i) {
            if (i < nt) {
                const TgtRec &r = rb[i];
                if (g0 < r.ext[0] && g1 < r.ext[1] && g2 < r.ext[2]) {
                    // W = -sumcell(...): the coefficients carry -M(j,k)/h
                    T w = 0;
                    if (r.nterm > 0) {
#pragma unroll
                        for (int q = 0; q < NT; ++q) {
                            const Term tm = r.t[q];
                            const int o = tm.off + lbase;
                            w += (T)tm.coef * (lds[o + tm.stride] - lds[o]);
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
                            const int s = a.fslot[app];
                            u = u * (r.F[s] ? reinterpret_cast<const T *>(r.F[s])[gidx]
                                            : (T)r.F_s[s]);
                        }
                    }
                    a.U[(int64_t)r.comp * a.slab + gidx] = u;
                }
            }
        }
        __syncthreads();
        if (gq + 1 < ng) commit_src();
        if (gq + 2 < ng) commit_blob(grp[gq + 2], buf);
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
    if (a.ngroups > MAX_GROUPS || a.ncf > MAX_CFIELDS) return hipErrorInvalidValue;
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
