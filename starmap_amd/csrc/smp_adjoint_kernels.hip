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

__device__ __forceinline__ int wrapc(int i, int n, int bc, int ext) {
    if (bc == 0) {
        if (i < 0) i += n;
        if (i >= n) i -= n;
        return i;
    }
    return i < 0 ? 0 : (i >= ext ? ext - 1 : i);
}

__device__ __forceinline__ double expm1div_d(double x, double *dedx) {
    if (fabs(x) > 2e-4) {
        const double ex = exp(x);
        *dedx = (ex * (x - 1) + 1) / (x * x);
        return (ex - 1) / x;
    }
    *dedx = .5 + x / 3;
    return 1 + x * .5 + (x * x) / 6;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(256) void adj_update_kernel(const AdjArgs<T> a) {
    const int grp = blockIdx.y;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int i0 = (int)(t % a.E[0]);
    const int64_t r = t / a.E[0];
    const int i1 = (int)(r % a.E[1]);
    const int i2 = (int)(r / a.E[1]);
    const int q0 = a.grp0[grp], q1 = a.grp0[grp + 1];
    const SMP_GLOBAL AdjComp<T> *comp = glb(a.comp);
    const SMP_GLOBAL int32_t *list = glb(a.list);
    {
        const int jf = list[q0];
        if (i2 >= a.E[2] || i0 >= comp[jf].ext[0] || i1 >= comp[jf].ext[1] || i2 >= comp[jf].ext[2])
            return;
    }
    const int ii[3] = {i0, i1, i2};
    const int64_t p = i0 + (int64_t)a.pitch0 * (i1 + (int64_t)a.ext1 * i2);
    const double hd = a.half_dt;
    const SMP_GLOBAL AdjTerm *rows = glb(a.rows);
    double ga = 0, gs = 0;
    for (int q = q0; q < q1; ++q) {
        const int j = list[q];
        const SMP_GLOBAL AdjComp<T> &c = comp[j];
        // W of the forward step (x, y, z terms in the forward order)
        double w = 0;
        for (int m = c.r0; m < c.r1; ++m) {
            const SMP_GLOBAL AdjTerm &tm = rows[m];
            const int e = tm.axis;
            int hi[3] = {ii[0], ii[1], ii[2]}, lo[3] = {ii[0], ii[1], ii[2]};
            if (comp[tm.k].cell[e])
                lo[e] = ii[e] - 1;   // face target between cells i-1 and i
            else
                hi[e] = ii[e] + 1;   // cell target between faces i and i+1
            const int ext = comp[tm.k].ext[e];
            hi[e] = wrapc(hi[e], a.n[e], a.bc[e], ext);
            lo[e] = wrapc(lo[e], a.n[e], a.bc[e], ext);
            const SMP_GLOBAL T *Uk = glb(a.U) + (int64_t)tm.k * a.slab;
            const double vh = (double)Uk[hi[0] + (int64_t)a.pitch0 * (hi[1] + (int64_t)a.ext1 * hi[2])];
            const double vl = (double)Uk[lo[0] + (int64_t)a.pitch0 * (lo[1] + (int64_t)a.ext1 * lo[2])];
            w += tm.coef * (vh - vl);
        }
        const double sg = c.sig ? (double)glb(c.sig)[p] : c.sig_s;
        const double E = c.E ? (double)glb(c.E)[p] : c.E_s;
        double dedx;
        (void)expm1div_d(-sg * hd, &dedx);
        const double dE = -hd * dedx;
        const double wq = w + (c.Q ? (double)glb(c.Q)[p] : 0.0);
        const int64_t jp = (int64_t)j * a.slab + p;
        // states before each application (the forward rounds to T after each)
        double u[2];
        u[0] = (double)glb(a.U0)[jp];
        u[1] = (double)(T)((u[0] + (hd * (wq - sg * u[0])) * E) * c.F[a.fslot[0] ? 1 : 0]);
        double lam = (double)glb(a.lam)[jp];
        double mu = 0, g = 0;
        for (int app = a.napp - 1; app >= 0; --app) {
            const double l = c.F[a.fslot[app] ? 1 : 0] * lam;
            mu += hd * E * l;
            g += l * (-hd * u[app] * E + hd * (wq - sg * u[app]) * dE);
            lam = (1 - hd * sg * E) * l;
        }
        glb(a.lam)[jp] = (T)lam;
        glb(a.mu)[jp] = (T)mu;
        ga += g;
        if (j > 0) gs += g;
    }
    const int64_t gp = (int64_t)a.gcode[grp] * a.slab + p;
    glb(a.Ga)[gp] += ga;
    glb(a.Gs)[gp] += gs;
}

template <typename T>
__global__ __launch_bounds__(256) void adj_transpose_kernel(const AdjArgs<T> a) {
    const int k = glb(a.list)[blockIdx.y];
    const SMP_GLOBAL AdjComp<T> &c = glb(a.comp)[k];
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int x0 = (int)(t % a.E[0]);
    const int64_t r = t / a.E[0];
    const int x1 = (int)(r % a.E[1]);
    const int x2 = (int)(r / a.E[1]);
    if (x2 >= a.E[2] || x0 >= c.ext[0] || x1 >= c.ext[1] || x2 >= c.ext[2]) return;
    const int xx[3] = {x0, x1, x2};
    double acc = 0;
    for (int q = c.c0; q < c.c1; ++q) {
        const SMP_GLOBAL AdjTerm &tm = glb(a.cols)[q];
        const int e = tm.axis;
        const int n = a.n[e];
        const SMP_GLOBAL T *mj = glb(a.mu) + (int64_t)tm.k * a.slab;
        // target indices whose high / low sample is x, and whether they count
        int ih[3] = {xx[0], xx[1], xx[2]}, il[3] = {xx[0], xx[1], xx[2]};
        bool use_h, use_l;
        if (c.cell[e]) {
            // face target s reads cells s (high) and s-1 (low)
            ih[e] = xx[e];
            il[e] = xx[e] + 1;
            if (a.bc[e] == 0) {
                il[e] = il[e] == n ? 0 : il[e];
                use_h = use_l = true;
            } else {
                // boundary faces 0 and n have zero differences
                use_h = ih[e] >= 1 && ih[e] <= n - 1;
                use_l = il[e] >= 1 && il[e] <= n - 1;
            }
        } else {
            // cell target i reads faces i+1 (high) and i (low)
            ih[e] = xx[e] - 1;
            il[e] = xx[e];
            if (a.bc[e] == 0) {
                ih[e] = ih[e] < 0 ? n - 1 : ih[e];
                use_h = use_l = true;
            } else {
                use_h = ih[e] >= 0;
                use_l = il[e] <= n - 1;
            }
        }
        double d = 0;
        if (use_h) d += (double)mj[ih[0] + (int64_t)a.pitch0 * (ih[1] + (int64_t)a.ext1 * ih[2])];
        if (use_l) d -= (double)mj[il[0] + (int64_t)a.pitch0 * (il[1] + (int64_t)a.ext1 * il[2])];
        acc += tm.coef * d;
    }
    const int64_t kp = (int64_t)k * a.slab + x0 + (int64_t)a.pitch0 * (x1 + (int64_t)a.ext1 * x2);
    glb(a.lam)[kp] = (T)((double)glb(a.lam)[kp] + acc);
}

template <typename T>
__global__ void copy_kernel(T *dst, const T *src, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T>
static dim3 adj_grid(const AdjArgs<T> &a, int y) {
    const int64_t cells = (int64_t)a.E[0] * a.E[1] * a.E[2];
    return dim3((unsigned)((cells + 255) / 256), (unsigned)y);
}

template <typename T>
hipError_t launch_adj_update(const AdjArgs<T> &a, hipStream_t s) {
    if (a.n_list <= 0) return hipSuccess;
    if (a.napp != 1 && a.napp != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_update_kernel<T>, adj_grid(a, a.n_grp), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adj_transpose(const AdjArgs<T> &a, hipStream_t s) {
    if (a.n_list <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_transpose_kernel<T>, adj_grid(a, a.n_list), dim3(256), 0, s, a);
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
    template hipError_t launch_adj_update<T>(const AdjArgs<T> &, hipStream_t);     \
    template hipError_t launch_adj_transpose<T>(const AdjArgs<T> &, hipStream_t);  \
    template hipError_t launch_copy<T>(T *, const T *, int64_t, hipStream_t);

SMP_ADJ_INSTANTIATE(double)
SMP_ADJ_INSTANTIATE(float)

}  // namespace smp
