// gfx950 kernels of the discrete adjoint of the StaRMAP time step.
//
// The forward half-step of target j at a point (starmap_solver.m:352-354,
// :404-406, with the substep filter of :355-358) is
//     u' = f * (u + a*(W + Q - s*u)*E),   a = dt/2,  E = expm1div(-s*a),
//     W  = sum_k coef_jk * (U_k(hi) - U_k(lo)),
// affine in u and W for fixed sigma s.  Its adjoint, for lambda' the
// adjoint of u':
//     l   = f * lambda'
//     mu += a*E * l                               (adjoint of W)
//     g  += l * (-a*u*E + a*(W + Q - s*u)*dE/ds)  (dJ/ds)
//     lambda = (1 - a*s*E) * l
// and the flux adjoint lambda_k += sum_j coef_jk * D^T mu_j is a gather over
// the transposed stencil.  dE/ds = -a * d/dx expm1div(x) at x = -s*a keeps
// the two branches of expm1div (starmap_solver.m:531-535).
#include "smp_kernels.h"

#include <cmath>

namespace smp {

#define SMP_GLOBAL __attribute__((address_space(1)))
template <typename P>
__device__ __forceinline__ SMP_GLOBAL P *glb(P *p) {
    return (SMP_GLOBAL P *)p;
}

namespace {

__device__ __forceinline__ double expm1div_d(double x, double *dedx) {
    if (fabs(x) > 2e-4) {
        const double ex = exp(x);
        *dedx = (ex * (x - 1) + 1) / (x * x);
        return (ex - 1) / x;
    }
    *dedx = .5 + x / 3;
    return 1 + x * .5 + (x * x) / 6;
}

// index i+d on an axis: wrapped on periodic axes, clamped into [0, ext)
// on extrapolation axes
__device__ __forceinline__ int adj_index(int i, int d, int n, int bc, int ext) {
    int j = i + d;
    if (bc == 0) return j < 0 ? j + n : (j >= n ? j - n : j);
    return j < 0 ? 0 : (j >= ext ? ext - 1 : j);
}

// The slab of component k as a buffer resource (base in scalar registers,
// num_records the slab's bytes): a load takes a 32-bit byte offset per lane
// and no 64-bit address arithmetic, and an offset past the slab (ADJ_OOB)
// reads zero.  Slabs are below 2 GiB (checked by the host).
constexpr uint32_t ADJ_OOB = 0xFFFFFF00u;
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab_rsrc(const T *base, int64_t slab_bytes, int k) {
    const char *b = reinterpret_cast<const char *>(base) + k * slab_bytes;
    // raw 32-bit loads: data format 32 in the descriptor's last dword
    return __builtin_amdgcn_make_buffer_rsrc((void *)b, (short)0, (int)slab_bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
}

// element at byte offset `off` of the slab of component k
template <typename T>
__device__ __forceinline__ T slab_at(const T *base, int64_t slab_bytes, int k, uint32_t off) {
    return buf_ld<T>(slab_rsrc(base, slab_bytes, k), off);
}

// Offsets (elements) of the high and low samples of the forward flux of a
// target at index i of an axis: a cell target (tcell) reads faces i+1 and
// i, a face target reads cells i and i-1 (starmap_solver.m:352-354 and the
// boundary extension of :359-377)
__device__ __forceinline__ void row_offsets(int i, bool tcell, int n, int bc, int stride, int &oh,
                                            int &ol) {
    const int ext = tcell ? n + bc : n;
    const int dh = tcell ? 1 : 0, dl = tcell ? 0 : -1;
    oh = (adj_index(i, dh, n, bc, ext) - i) * stride;
    ol = (adj_index(i, dl, n, bc, ext) - i) * stride;
}

// Transposed: a cell component at index i is the high sample of face
// target i and the low sample of face target i+1 (targets in [1, n) count
// on extrapolation axes, the boundary faces carry no difference); a face
// component is the high sample of cell target i-1 and the low sample of
// cell target i
__device__ __forceinline__ void col_offsets(int i, bool kcell, int n, int bc, int stride, int &oh,
                                            int &ol, bool &mh, bool &ml) {
    const int ext = kcell ? n + bc : n;
    const int dh = kcell ? 0 : -1, dl = kcell ? 1 : 0, lb = kcell ? 1 : 0;
    const int ih = i + dh, il = i + dl;
    mh = bc == 0 || (ih >= lb && ih < n);
    ml = bc == 0 || (il >= lb && il < n);
    oh = (adj_index(i, dh, n, bc, ext) - i) * stride;
    ol = (adj_index(i, dl, n, bc, ext) - i) * stride;
}

// Point rows of a thread: 64 consecutive i0 per wave, the row index (i1, i2)
// the same across the wave (made explicit for the compiler)
__device__ __forceinline__ int wave_row(int row) { return __builtin_amdgcn_readfirstlane(row); }

// Sample offsets of a point on each kernel axis (per lane on axis 0, the
// same across the wave on axes 1 and 2) and, for transposed terms, whether
// each sample counts.
struct AxisOffsets {
    int h[3], l[3];
    bool mh[3], ml[3];
};

// One group of terms: coefficients and the high and low samples at element
// pe + offsets of the group's axis.
template <typename T>
struct Group {
    T vh[ADJ_GROUP], vl[ADJ_GROUP];
    double coef[ADJ_GROUP];
};

// the samples of a group of terms; a sample that does not count (mh / ml
// false, transposed terms) reads zero through an offset past the slab
template <typename T>
__device__ __forceinline__ void gather(Group<T> &gr, const T *base, int64_t sb,
                                       const SMP_GLOBAL AdjTerm *g, int pe, int oh, int ol, bool mh,
                                       bool ml) {
    int k[ADJ_GROUP];
#pragma unroll
    for (int t = 0; t < ADJ_GROUP; ++t) {
        k[t] = g[t].k;
        gr.coef[t] = g[t].coef;
    }
    const uint32_t bh = mh ? (uint32_t)(pe + oh) * (uint32_t)sizeof(T) : ADJ_OOB;
    const uint32_t bl = ml ? (uint32_t)(pe + ol) * (uint32_t)sizeof(T) : ADJ_OOB;
#pragma unroll
    for (int t = 0; t < ADJ_GROUP; ++t) {
        const __amdgpu_buffer_rsrc_t r = slab_rsrc(base, sb, k[t]);
        gr.vh[t] = buf_ld<T>(r, bh);
        gr.vl[t] = buf_ld<T>(r, bl);
    }
}

template <typename T>
__device__ __forceinline__ double reduce(const Group<T> &gr) {
    // the difference in the state precision (a sample that does not count
    // was read as zero), the sum in FP64
    double w = 0;
#pragma unroll
    for (int t = 0; t < ADJ_GROUP; ++t) {
        const T d = gr.vh[t] - gr.vl[t];
        w += gr.coef[t] * (double)d;
    }
    return w;
}

// sum over a component's term list of coef*(high - low) (masked for
// transposed terms), axes in order: the first group of every axis gathered
// by gather_first (the caller issues every first group before reducing any)
template <typename T, int NDIM>
__device__ __forceinline__ void gather_first(Group<T> *g0, const T *base, int64_t sb,
                                             const SMP_GLOBAL AdjTerm *terms, const int32_t *ra, int pe,
                                             const AxisOffsets &o) {
#pragma unroll
    for (int e = 0; e < NDIM; ++e) gather(g0[e], base, sb, terms + ra[e], pe, o.h[e], o.l[e], o.mh[e], o.ml[e]);
}

template <typename T, int NDIM>
__device__ __forceinline__ double flux_rest(const Group<T> *g0, const T *base, int64_t sb,
                                            const SMP_GLOBAL AdjTerm *terms, const int32_t *ra, int pe,
                                            const AxisOffsets &o) {
    double w = 0;
#pragma unroll
    for (int e = 0; e < NDIM; ++e) {
        w += reduce(g0[e]);
        for (int m = ra[e] + ADJ_GROUP; m < ra[e + 1]; m += ADJ_GROUP) {
            Group<T> gr;
            gather(gr, base, sb, terms + m, pe, o.h[e], o.l[e], o.mh[e], o.ml[e]);
            w += reduce(gr);
        }
    }
    return w;
}

}  // namespace

// Adjoint of one launch for the targets of one grid-type group per item.
// FUSE: first lambda += the transposed flux of the launch after it (its mu,
// the other parity, terms by source column), as adj_transpose_kernel stores
// it.
// HASW: W+Q of every target comes from the forward launch's record (a.W);
// otherwise the kernel gathers the forward flux itself.
template <typename T, bool FUSE, int NDIM, bool HASW>
__global__ __launch_bounds__(256) void adj_update_kernel(const AdjArgs<T> a) {
    // Work order: grid-type groups of the same points next to each other
    // (they share source components), then 64-point chunks along i0, then
    // groups of four rows; the hardware deals workgroups to the 8 XCDs in
    // turn, so workgroup b takes the item that gives each XCD (and its L2)
    // one contiguous band of rows
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int xcd = b & 7, per = nb >> 3, rem = nb & 7;
    const int item = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
    const int grp = item % a.n_grp;
    const int nchunk = (a.E[0] + 63) >> 6;
    const int rest = item / a.n_grp;
    const int row = wave_row((rest / nchunk) * 4 + (int)threadIdx.y);
    if (row >= a.E[1] * a.E[2]) return;
    const int q0 = a.grp0[grp], q1 = a.grp0[grp + 1];
    const SMP_GLOBAL AdjComp<T> *comp = glb(a.comp);
    const SMP_GLOBAL int32_t *list = glb(a.list);
    const int jf = list[q0];
    int i[3];
    i[0] = (rest % nchunk) * 64 + (int)threadIdx.x;
    i[1] = row % a.E[1];
    i[2] = row / a.E[1];
    if (i[0] >= comp[jf].ext[0] || i[1] >= comp[jf].ext[1] || i[2] >= comp[jf].ext[2]) return;
    const int pe = i[0] + a.pitch0 * (i[1] + a.ext1 * i[2]);
    // sample offsets of the group's grid type
    const int stride[3] = {1, a.pitch0, a.pitch0 * a.ext1};
    AxisOffsets ro, co;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const bool cl = comp[jf].cell[e] != 0;
        row_offsets(i[e], cl, a.n[e], a.bc[e], stride[e], ro.h[e], ro.l[e]);
        ro.mh[e] = ro.ml[e] = true;
        if (FUSE) col_offsets(i[e], cl, a.n[e], a.bc[e], stride[e], co.h[e], co.l[e], co.mh[e], co.ml[e]);
    }
    const double hd = a.half_dt;
    const int64_t sb = a.slab * (int64_t)sizeof(T);
    const uint32_t pb = (uint32_t)pe * sizeof(T);
    double ga = 0, gs = 0;
    for (int q = q0; q < q1; ++q) {
        const int j = list[q];
        const SMP_GLOBAL AdjComp<T> &c = comp[j];
        int32_t ra[4], ca[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            ra[e] = c.ra[e];
            ca[e] = c.ca[e];
        }
        // every load of the component (the first group of each term list
        // on each axis included) issued before the first use: fields that
        // are scalars read U0 instead
        const bool has_sig = c.sig != nullptr, has_E = c.E != nullptr, has_Q = c.Q != nullptr;
        const T u0_t = slab_at(a.U0, sb, j, pb);
        const T lam_t = slab_at(a.lam, sb, j, pb);
        const T sig_t = slab_at(has_sig ? c.sig : a.U0, sb, 0, pb);
        const bool has_dE = c.dE != nullptr;
        const T E_t = slab_at(has_E ? c.E : a.U0, sb, 0, pb);
        const T dE_t = slab_at(has_dE ? c.dE : a.U0, sb, 0, pb);
        const T Q_t = slab_at(has_Q ? c.Q : a.U0, sb, 0, pb);
        const T W_t = HASW ? slab_at(a.W, sb, j, pb) : (T)0;
        Group<T> gr[NDIM], gc[NDIM];
        if (!HASW) gather_first<T, NDIM>(gr, a.U, sb, glb(a.rows), ra, pe, ro);
        if (FUSE) gather_first<T, NDIM>(gc, a.mu, sb, glb(a.cols), ca, pe, co);
        // W of the forward step (terms in the forward order)
        const double w = HASW ? 0.0 : flux_rest<T, NDIM>(gr, a.U, sb, glb(a.rows), ra, pe, ro);
        double lam = (double)lam_t;
        // lambda of the launch after this one, as the transposed-flux
        // kernel stores it
        if (FUSE) lam = (double)(T)(lam + flux_rest<T, NDIM>(gc, a.mu, sb, glb(a.cols), ca, pe, co));
        const double sg = has_sig ? (double)sig_t : c.sig_s;
        const double E = has_E ? (double)E_t : c.E_s;
        const double wq = HASW ? (double)W_t : w + (has_Q ? (double)Q_t : 0.0);
        const double dE = has_dE ? (double)dE_t : c.dE_s;
        // states before each application (the forward rounds to T after each)
        double u[2];
        u[0] = (double)u0_t;
        u[1] = (double)(T)((u[0] + (hd * (wq - sg * u[0])) * E) * c.F[a.fslot[0] ? 1 : 0]);
        double mu = 0, g = 0;
        for (int app = a.napp - 1; app >= 0; --app) {
            const double l = c.F[a.fslot[app] ? 1 : 0] * lam;
            mu += hd * E * l;
            g += l * (-hd * u[app] * E + hd * (wq - sg * u[app]) * dE);
            lam = (1 - hd * sg * E) * l;
        }
        const int64_t jp = (int64_t)j * a.slab + pe;
        glb(a.lam)[jp] = (T)lam;
        glb(a.mu)[jp] = (T)mu;
        ga += g;
        if (j > 0) gs += g;
    }
    const int64_t gp = (int64_t)a.gcode[grp] * a.slab + pe;
    glb(a.Ga)[gp] += ga;
    glb(a.Gs)[gp] += gs;
}

// lambda_k += the transposed flux of mu for the components k in a.list
// (grid.z), ADJ_ROWS rows of points per thread
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void adj_transpose_kernel(const AdjArgs<T> a) {
    const int k = glb(a.list)[blockIdx.z];
    const SMP_GLOBAL AdjComp<T> &c = glb(a.comp)[k];
    const int ext0 = c.ext[0], ext1 = c.ext[1], ext2 = c.ext[2];
    const int i0 = (int)blockIdx.y * 64 + (int)threadIdx.x;
    if (i0 >= ext0) return;
    int32_t ca[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) ca[e] = c.ca[e];
    const int64_t sb = a.slab * (int64_t)sizeof(T);
    const int stride[3] = {1, a.pitch0, a.pitch0 * a.ext1};
    // rows outside the component read its last row and store nothing
    const int nrows = a.E[1] * a.E[2];
    const int row0 = wave_row(((int)blockIdx.x * 4 + (int)threadIdx.y) * ADJ_ROWS);
    int pe[ADJ_ROWS];
    bool live[ADJ_ROWS];
    AxisOffsets co[ADJ_ROWS];
    T lam0[ADJ_ROWS];
#pragma unroll
    for (int r = 0; r < ADJ_ROWS; ++r) {
        const int row = row0 + r;
        int i[3];
        i[0] = i0;
        i[1] = row % a.E[1];
        i[2] = row / a.E[1];
        live[r] = row < nrows && i[1] < ext1 && i[2] < ext2;
        i[1] = i[1] < ext1 ? i[1] : ext1 - 1;
        i[2] = i[2] < ext2 ? i[2] : ext2 - 1;
        pe[r] = i[0] + a.pitch0 * (i[1] + a.ext1 * i[2]);
#pragma unroll
        for (int e = 0; e < 3; ++e)
            col_offsets(i[e], c.cell[e] != 0, a.n[e], a.bc[e], stride[e], co[r].h[e], co[r].l[e],
                        co[r].mh[e], co[r].ml[e]);
        lam0[r] = slab_at(a.lam, sb, k, (uint32_t)pe[r] * sizeof(T));
    }
    Group<T> g0[ADJ_ROWS][NDIM];
#pragma unroll
    for (int r = 0; r < ADJ_ROWS; ++r) gather_first<T, NDIM>(g0[r], a.mu, sb, glb(a.cols), ca, pe[r], co[r]);
    double acc[ADJ_ROWS];
#pragma unroll
    for (int r = 0; r < ADJ_ROWS; ++r)
        acc[r] = flux_rest<T, NDIM>(g0[r], a.mu, sb, glb(a.cols), ca, pe[r], co[r]);
#pragma unroll
    for (int r = 0; r < ADJ_ROWS; ++r)
        if (live[r]) glb(a.lam)[(int64_t)k * a.slab + pe[r]] = (T)((double)lam0[r] + acc[r]);
}

// dE/dsigma of E = expm1div(-sigma*dt/2): -dt/2 * expm1div'(-sigma*dt/2)
template <typename T>
__global__ void dedsig_kernel(const T *sig, T *dE, int64_t count, double dt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        double dedx;
        (void)expm1div_d(-(double)glb(sig)[i] * (dt / 2), &dedx);
        glb(dE)[i] = (T)(-(dt / 2) * dedx);
    }
}

template <typename T>
__global__ void copy_kernel(T *dst, const T *src, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T>
static dim3 adj_grid(const AdjArgs<T> &a, int z, int rows_per_thread) {
    const int64_t rows = (int64_t)a.E[1] * a.E[2];
    const int64_t per_block = 4 * rows_per_thread;
    return dim3((unsigned)((rows + per_block - 1) / per_block), (unsigned)((a.E[0] + 63) / 64),
                (unsigned)z);
}

template <typename T>
hipError_t launch_adj_update(const AdjArgs<T> &a, bool fuse, hipStream_t s) {
    if (a.n_list <= 0) return hipSuccess;
    if (a.napp != 1 && a.napp != 2) return hipErrorInvalidValue;
    const int64_t rows = (int64_t)a.E[1] * a.E[2];
    const int64_t items = (rows + 3) / 4 * ((a.E[0] + 63) / 64) * a.n_grp;
    if (items >= (int64_t)1 << 31) return hipErrorInvalidValue;
    const dim3 grid((unsigned)items), block(64, 4);
#define SMP_ADJ_UPD(F, D)                                                                \
    do {                                                                                 \
        if (a.W)                                                                         \
            hipLaunchKernelGGL((adj_update_kernel<T, F, D, true>), grid, block, 0, s, a);  \
        else                                                                             \
            hipLaunchKernelGGL((adj_update_kernel<T, F, D, false>), grid, block, 0, s, a); \
    } while (0)
    switch (a.ndim * 2 + (fuse ? 1 : 0)) {
    case 2: SMP_ADJ_UPD(false, 1); break;
    case 3: SMP_ADJ_UPD(true, 1); break;
    case 4: SMP_ADJ_UPD(false, 2); break;
    case 5: SMP_ADJ_UPD(true, 2); break;
    case 6: SMP_ADJ_UPD(false, 3); break;
    case 7: SMP_ADJ_UPD(true, 3); break;
    default: return hipErrorInvalidValue;
    }
#undef SMP_ADJ_UPD
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adj_transpose(const AdjArgs<T> &a, hipStream_t s) {
    if (a.n_list <= 0) return hipSuccess;
    const dim3 grid = adj_grid(a, a.n_list, ADJ_ROWS), block(64, 4);
    switch (a.ndim) {
    case 1: hipLaunchKernelGGL((adj_transpose_kernel<T, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((adj_transpose_kernel<T, 2>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((adj_transpose_kernel<T, 3>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dedsig(const T *sig, T *dE, int64_t count, double dt, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(dedsig_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, sig, dE, count, dt);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(T *dst, const T *src, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(copy_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, count);
    return hipGetLastError();
}

#define SMP_ADJ_INSTANTIATE(T)                                                     \
    template hipError_t launch_adj_update<T>(const AdjArgs<T> &, bool, hipStream_t); \
    template hipError_t launch_adj_transpose<T>(const AdjArgs<T> &, hipStream_t);  \
    template hipError_t launch_copy<T>(T *, const T *, int64_t, hipStream_t);    \
    template hipError_t launch_dedsig<T>(const T *, T *, int64_t, double, hipStream_t);

SMP_ADJ_INSTANTIATE(double)
SMP_ADJ_INSTANTIATE(float)

}  // namespace smp
