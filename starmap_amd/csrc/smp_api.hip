// C ABI of the StaRMAP time-stepping core (include/starmap_hip.h).
//
// Device layout (moment-major SoA): component j occupies one slab of
// `slab` elements; inside a slab, cell (i0,i1,i2) over the kernel axes sits
// at i0 + pitch0*(i1 + E1*i2), with E_e = n_e + bc_e so that the n+1 faces
// of an extrapolation axis fit, and pitch0 padded to 16 elements.  Kernel
// axes are the solver axes that carry cells or couplings (2D: solver y,z).
// On a periodic axis, face f of the reference (at ax_lo+(f+1)h) is stored at
// index (f+1) mod n, so that on every axis face s sits at ax_lo + s*h and
// the difference to cell centres is U(s)-U(s-1) / F(c+1)-F(c).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/starmap_hip.h"
#include "smp_kernels.h"
#include "smp_plan.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// On failure the runtime's last-error state is cleared as well, so that a
// later launch check (hipGetLastError) does not report this error again.
#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                      \
            return fail(e_ == hipErrorOutOfMemory ? SMP_E_OOM : SMP_E_HIP,                \
                        std::string(#expr) + ": " + hipGetErrorString(e_));               \
        }                                                                                 \
    } while (0)

#define REQUIRE_DEVICE(h)                                                                 \
    do {                                                                                  \
        if ((h)->host_only)                                                               \
            return fail(SMP_E_STATE, "handle was created without a device (device = -1)"); \
    } while (0)

double expm1div_host(double x) {
    // starmap_solver.m:531-535
    if (std::fabs(x) > 2e-4) return (std::exp(x) - 1) / x;
    return 1 + x * .5 + (x * x) / 6;
}

struct Field {
    void *dev = nullptr;     // device field (slab layout) or null
    double scalar = 0.0;
    bool set = false;
};

using Key = std::tuple<int, int, int, int, int>;   // gt x,y,z, order, slot

}  // namespace

struct smp_problem {
    smp::Geometry geo;
    smp::TileShape ts;
    smp::Plan plan[2];
    int precision = SMP_FP64;
    int elem = 8;
    int f_position = SMP_FILTER_NONE;
    int device = 0;
    hipStream_t stream = nullptr;
    int E[3] = {1, 1, 1};          // storage extent per kernel axis
    int pitch0 = 1;
    int64_t slab = 0;
    int ntile[3] = {1, 1, 1};
    int n_tiles = 0;
    void *U = nullptr;
    void *plan_dev[2] = {nullptr, nullptr};
    size_t plan_off[2][3] = {{0, 0, 0}, {0, 0, 0}};   // loads, targets offsets
    size_t lds_bytes[2] = {0, 0};
    int ncf_cap[2] = {0, 0};       // coefficient fields staged per tile (LDS reserved)
    int ncf[2] = {0, 0};           // coefficient fields currently staged
    bool gen[2] = {false, false};  // some coefficient is read from global memory
    void *cf_dev[2] = {nullptr, nullptr};
    int blocks_per_cu = 0;
    int cell_waves = 4;            // 4, or 8 (two waves per SIMD on every cell)
    bool host_only = false;
    std::vector<int32_t> mom_order, f_order;
    Field sA1, EA;
    std::map<Key, Field> sT, ET, F;
    std::vector<Field> Q;
    double dt_decay = std::nan("");
    bool decay_dirty = true;
    bool coef_dirty = true;
    std::vector<int32_t> snap_comps;
    void *snap = nullptr;
    void *interp_buf = nullptr;
    int64_t interp_cap = 0;
    void *red_buf = nullptr;      // reduction partials, results and the non-finite flag
    int check_every = 64;         // steps between non-finite scans in smp_step (0: none)
    int64_t steps_unchecked = 0;
    void *norm_buf = nullptr;   // partials and results of smp_moment_norms
    void *sink = nullptr;         // 1024 elements: stores of lanes without a target cell
    std::vector<hipEvent_t> events;
    std::vector<void *> ckpt;     // device copies of the whole state
    // discrete adjoint: lambda, mu (state layout), dJ/dsigma per component
    // (FP64), tables of the adjoint kernels, segment checkpoints and the
    // states recorded while a segment is recomputed
    void *lam = nullptr, *mu = nullptr, *G = nullptr;
    hipStream_t adj_stream = nullptr;   // the adjoint's recomputation (default priority)
    // the adjoint's events, created once: timing (forward start, reverse
    // start, end) and ordering between the two streams (forward done,
    // recomputation and reverse sweep of each record set)
    hipEvent_t adj_ev[8] = {};
    int adj_ngrp[2] = {0, 0};     // odd, even: grid types, list ranges, codes
    int32_t adj_grp0[2][9] = {};
    int32_t adj_gcode[2][8] = {};
    void *adj_comp = nullptr, *adj_rows = nullptr, *adj_cols = nullptr, *adj_list = nullptr;
    // adjoint schedule buffers: checkpoint slots, record ranges (each of
    // adj_range_rb state-sized buffers, two launches' records per buffer),
    // the scratch state of re-advanced segments and the tail's start state
    std::vector<void *> adj_slot, adj_range;
    int64_t adj_range_rb = 0;
    void *adj_work = nullptr, *adj_tstart = nullptr;
    std::vector<hipEvent_t> adj_range_ev;   // per range: recomputed, reversed
    int64_t adj_budget = 0;                 // bytes (0: one checkpoint per segment)
    smp::AdjSchedule adj_sched;
    int64_t adj_peak_bytes = 0;
    // the LDS-staged adjoint update: its tile, the plans of both parities,
    // their device records, load blobs and staged coefficient fields
    bool adj_lds_ok = false;
    smp::TileShape adj_ts;
    int adj_ntile[3] = {1, 1, 1};
    int adj_ntiles = 0;
    smp::AdjPlan adj_plan[2];
    void *adj_lds_recs[2] = {nullptr, nullptr}, *adj_lds_blobs[2] = {nullptr, nullptr};
    void *adj_lds_cf[2] = {nullptr, nullptr}, *adj_lds_gf[2] = {nullptr, nullptr};
    void *adj_zeros = nullptr;   // 64 zero dwords: boundary faces of masked mu tiles
    int adj_lds_ncf[2] = {0, 0};
    bool adj_recw = false;   // the recording launches store W+Q (set per smp_adjoint call)
    bool adj_lds_gen[2] = {false, false};   // some coefficient field is read from global memory
    size_t adj_lds_bytes[2] = {0, 0};
    std::map<const void *, void *> adj_dE;   // dE/dsigma field of each sigma field
    double adj_fwd_ms = 0, adj_rev_ms = 0;
    double last_total_ms = 0, last_substep_ms = 0, last_bytes = 0;
    int64_t last_launches = 0;
    int64_t dev_bytes = 0;
    std::vector<void *> allocs;

    ~smp_problem() {
        if (adj_stream) (void)hipStreamSynchronize(adj_stream);
        for (void *p : allocs) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipEvent_t e : adj_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : adj_range_ev) (void)hipEventDestroy(e);
        if (adj_stream) (void)hipStreamDestroy(adj_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int dalloc(void **p, size_t bytes) {
        hipError_t e = hipMalloc(p, bytes ? bytes : 1);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(e == hipErrorOutOfMemory ? SMP_E_OOM : SMP_E_HIP,
                        std::string("hipMalloc(") + std::to_string(bytes) + "): " +
                            hipGetErrorString(e));
        }
        allocs.push_back(*p);
        dev_bytes += (int64_t)bytes;
        return SMP_OK;
    }

    void dfree(void *p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree(p);
    }

    int64_t numel(int j) const {
        int64_t m = 1;
        for (int d = 0; d < 3; ++d) m *= geo.ext_of(j, d);
        return m;
    }

    int64_t numel_gt(const int gt[3]) const {
        int64_t m = 1;
        for (int d = 0; d < 3; ++d) m *= gt[d] == 1 ? geo.n[d] : geo.n[d] + geo.bc[d];
        return m;
    }

    // Convert between the reference layout of a field on grid gt (FP64,
    // column-major over solver axes) and the slab layout of type T.
    template <typename T>
    void pack(const int gt[3], const double *src, T *dst) const {
        int e[3];
        for (int d = 0; d < 3; ++d) e[d] = gt[d] == 1 ? geo.n[d] : geo.n[d] + geo.bc[d];
        // per solver axis: map index -> (kernel axis stride contribution)
        std::vector<int64_t> off[3];
        for (int d = 0; d < 3; ++d) {
            off[d].resize(e[d]);
            int ka = -1;
            for (int k = 0; k < geo.ndim; ++k)
                if (geo.axis[k] == d) ka = k;
            int64_t stride = ka == 0 ? 1 : ka == 1 ? pitch0 : (int64_t)pitch0 * E[1];
            for (int i = 0; i < e[d]; ++i) {
                int s = i;
                if (geo.bc[d] == 0 && gt[d] == 2) s = (i + 1) % geo.n[d];
                off[d][i] = ka < 0 ? 0 : s * stride;
            }
        }
        std::fill(dst, dst + slab, (T)0);
        int64_t p = 0;
        for (int i3 = 0; i3 < e[2]; ++i3)
            for (int i2 = 0; i2 < e[1]; ++i2) {
                const int64_t b = off[2][i3] + off[1][i2];
                for (int i1 = 0; i1 < e[0]; ++i1) dst[b + off[0][i1]] = (T)src[p++];
            }
    }

    template <typename T>
    void unpack(const int gt[3], const T *src, double *dst) const {
        int e[3];
        for (int d = 0; d < 3; ++d) e[d] = gt[d] == 1 ? geo.n[d] : geo.n[d] + geo.bc[d];
        std::vector<int64_t> off[3];
        for (int d = 0; d < 3; ++d) {
            off[d].resize(e[d]);
            int ka = -1;
            for (int k = 0; k < geo.ndim; ++k)
                if (geo.axis[k] == d) ka = k;
            int64_t stride = ka == 0 ? 1 : ka == 1 ? pitch0 : (int64_t)pitch0 * E[1];
            for (int i = 0; i < e[d]; ++i) {
                int s = i;
                if (geo.bc[d] == 0 && gt[d] == 2) s = (i + 1) % geo.n[d];
                off[d][i] = ka < 0 ? 0 : s * stride;
            }
        }
        int64_t p = 0;
        for (int i3 = 0; i3 < e[2]; ++i3)
            for (int i2 = 0; i2 < e[1]; ++i2) {
                const int64_t b = off[2][i3] + off[1][i2];
                for (int i1 = 0; i1 < e[0]; ++i1) dst[p++] = (double)src[b + off[0][i1]];
            }
    }

    // upload a host FP64 field on grid gt into a (new or existing) slab
    int upload_field(const int gt[3], const double *src, void **dev) {
        if (!*dev) {
            int rc = dalloc(dev, (size_t)slab * elem);
            if (rc) return rc;
        }
        if (precision == SMP_FP64) {
            std::vector<double> tmp(slab);
            pack<double>(gt, src, tmp.data());
            HIP_TRY(hipMemcpy(*dev, tmp.data(), (size_t)slab * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> tmp(slab);
            pack<float>(gt, src, tmp.data());
            HIP_TRY(hipMemcpy(*dev, tmp.data(), (size_t)slab * 4, hipMemcpyHostToDevice));
        }
        return SMP_OK;
    }
};

namespace {

// Build the schedule of one substep kind for a reserve of staged coefficient
// fields.  Workgroups per CU (SMP_BLOCKS_PER_CU, default 1): fewer LDS slots
// per workgroup trade source re-staging (mostly L2 / Infinity Cache hits) for
// occupancy; when the slots are too few for the coupling the schedule falls
// back to fewer workgroups.  Within the chosen budget the schedule shrinks to
// the smallest slot count that keeps the group count and stages every source
// once, when that fits.
int make_plan(smp_problem *h, int kind, int ncf_reserve) {
    std::string err = "no schedule";
    const size_t slot_bytes = (size_t)h->ts.slot_elems * h->elem;
    const size_t fixed = smp::substep_lds_overhead() + (size_t)ncf_reserve * h->ts.threads * h->elem;
    const int max_load_elems = smp::MAXL * h->ts.slot_elems;
    smp::Plan plan;
    int blocks = h->blocks_per_cu;
    int cap = 0;
    for (; blocks >= 1; --blocks) {
        // 256 bytes of the 160 KB stay free (static LDS of the batched kernel)
        const size_t per_block = (160 * 1024 - 256) / blocks;
        if (per_block <= fixed + 2 * slot_bytes) continue;
        cap = (int)((per_block - fixed) / slot_bytes);
        err = smp::build_plan(h->geo, h->ts, kind, cap, smp::GMAX, max_load_elems, smp::MAXL, plan);
        if (err.empty()) break;
    }
    if (!err.empty()) return fail(SMP_E_INVALID, err);
    if ((int)plan.loads.size() == plan.n_sources) {
        // fewest slots with the same group count (every group is one barrier
        // phase) and no re-staging
        const size_t ngroups = plan.groups.size();
        int lo = 2, hi = cap;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            smp::Plan trial;
            std::string e2 = smp::build_plan(h->geo, h->ts, kind, mid, smp::GMAX, max_load_elems,
                                             smp::MAXL, trial);
            if (e2.empty() && (int)trial.loads.size() == trial.n_sources &&
                trial.groups.size() <= ngroups)
                hi = mid;
            else
                lo = mid + 1;
        }
        err = smp::build_plan(h->geo, h->ts, kind, lo, smp::GMAX, max_load_elems, smp::MAXL, plan);
        if (!err.empty()) return fail(SMP_E_INVALID, err);
    }
    h->plan[kind] = plan;
    h->ncf_cap[kind] = ncf_reserve;
    h->lds_bytes[kind] = fixed + (size_t)std::max(plan.n_slots, 1) * slot_bytes;
    if (h->host_only) return SMP_OK;
    // per-group record blobs, filled in whenever coefficients change
    h->dfree(h->plan_dev[kind]);
    h->plan_dev[kind] = nullptr;
    return h->dalloc(&h->plan_dev[kind], std::max<size_t>(1, plan.groups.size()) * smp::BLOB_SZ);
}

// Scan the whole state for NaN or infinite values (one pass over the
// state; the flag lives after the reduction results in red_buf and is
// cleared by a fill kernel).
int scan_state(smp_problem *h, bool *bad) {
    int *flag = (int *)((double *)h->red_buf + 3 * 1024 + 3);
    const int64_t count = h->slab * h->geo.n_sys;
    HIP_TRY(smp::launch_fill(flag, 1, 0, h->stream));
    if (h->precision == SMP_FP64)
        HIP_TRY(smp::launch_finite<double>((const double *)h->U, count, flag, h->stream));
    else
        HIP_TRY(smp::launch_finite<float>((const float *)h->U, count, flag, h->stream));
    int v = 0;
    HIP_TRY(hipMemcpyAsync(&v, flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->steps_unchecked = 0;
    *bad = v != 0;
    return SMP_OK;
}

// Per component, the stored indices and weights of the 2^ndim grid points
// around `point` on the component's own staggered grid (NaN outside).
void probe_table(const smp_problem *h, const double *lo, const double *point,
                 std::vector<smp::ProbeComp> &pc) {
    const smp::Geometry &g = h->geo;
    const int64_t stride[3] = {1, h->pitch0, (int64_t)h->pitch0 * h->E[1]};
    pc.assign(g.n_sys, smp::ProbeComp());
    for (int j = 0; j < g.n_sys; ++j) {
        smp::ProbeComp &c = pc[j];
        std::memset(&c, 0, sizeof(c));
        // per kernel axis: the two stored indices and weights of the point
        int64_t s0[3] = {0, 0, 0}, s1[3] = {0, 0, 0};
        double t[3] = {0, 0, 0};
        for (int e = 0; e < g.ndim; ++e) {
            const int ax = g.axis[e];
            const bool face = g.gt[3 * j + ax] == 2;
            const int n = g.n[ax], bc = g.bc[ax];
            const double hh = g.h[ax];
            // the reference's grid vector of this grid type (starmap_solver.m:113-118)
            const double first = lo[ax] + (face ? (bc ? 0.0 : hh) : hh / 2);
            const int len = face ? n + bc : n;
            double f = (point[ax] - first) / hh;
            // the grid vector is first + k*h here and MATLAB's colon there:
            // points within 1e-9 cells of either end are on it
            if (f < 0 && f > -1e-9) f = 0;
            if (f > len - 1 && f < len - 1 + 1e-9) f = len - 1;
            if (!(f >= 0) || f > len - 1) {
                c.nan = 1;
                break;
            }
            int f0 = (int)std::floor(f);
            if (f0 > len - 2) f0 = len > 1 ? len - 2 : 0;
            t[e] = len > 1 ? f - f0 : 0.0;
            const int f1 = len > 1 ? f0 + 1 : f0;
            auto stored = [&](int fi) { return face && bc == 0 ? (fi + 1) % n : fi; };
            s0[e] = stored(f0);
            s1[e] = stored(f1);
        }
        if (c.nan) continue;
        const int corners = 1 << g.ndim;
        for (int k = 0; k < corners; ++k) {
            double w = 1;
            int64_t off = 0;
            for (int e = 0; e < g.ndim; ++e) {
                const bool hi = (k >> e) & 1;
                w *= hi ? t[e] : 1 - t[e];
                off += (hi ? s1[e] : s0[e]) * stride[e];
            }
            c.off[k] = off;
            c.w[k] = w;
        }
        c.n = corners;
    }
}

// Upload the table to dpc, interpolate into dout and copy the values out.
int probe_run(smp_problem *h, const std::vector<smp::ProbeComp> &pc, smp::ProbeComp *dpc,
              double *dout, double *out) {
    const smp::Geometry &g = h->geo;
    HIP_TRY(hipMemcpyAsync(dpc, pc.data(), sizeof(smp::ProbeComp) * pc.size(), hipMemcpyHostToDevice,
                           h->stream));
    if (h->precision == SMP_FP64)
        HIP_TRY(smp::launch_probe<double>((const double *)h->U, h->slab, dpc, g.n_sys, dout, h->stream));
    else
        HIP_TRY(smp::launch_probe<float>((const float *)h->U, h->slab, dpc, g.n_sys, dout, h->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, sizeof(double) * pc.size(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

// The fixed device array of smp_probe_noalloc: the component table and the
// values of SMP_PROBE_NOALLOC_MAX_COMPS components (about 590 KB on every
// GPU the library runs on, held until the process ends).  Host code never
// uses the symbol itself, only the address hipGetSymbolAddress returns for
// the current device, resolved once per device when a handle is created.
static __device__ __align__(256) unsigned char
    scratch[SMP_PROBE_NOALLOC_MAX_COMPS * (sizeof(smp::ProbeComp) + sizeof(double))];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// Resolve the array's address on `device` (which must be current).
int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(SMP_E_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return SMP_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(SMP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = (char *)ptr;
    return SMP_OK;
}

int probe_scratch(int device, char **base) {
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices && g_scratch_ptr[device]) {
            *base = g_scratch_ptr[device];
            return SMP_OK;
        }
    }
    int rc = resolve_scratch(device);
    if (rc) return rc;
    *base = g_scratch_ptr[device];
    return SMP_OK;
}

}  // namespace

extern "C" {

const char *smp_last_error(void) { return g_err.c_str(); }

const char *smp_version(void) { return "starmap-hip 0.1 (gfx950)"; }

int smp_create(const smp_desc *d, smp_problem **out) {
    g_err.clear();
    if (!d || !out) return fail(SMP_E_INVALID, "null argument");
    *out = nullptr;
    if (d->precision != SMP_FP64 && d->precision != SMP_FP32)
        return fail(SMP_E_INVALID, "precision must be SMP_FP64 or SMP_FP32");
    if (d->f_position < SMP_FILTER_NONE || d->f_position > SMP_FILTER_FULL)
        return fail(SMP_E_INVALID, "invalid f_position");
    if (d->f_position != SMP_FILTER_NONE && !d->f_order)
        return fail(SMP_E_INVALID, "a filter needs f_order");
    if (!d->mom_order) return fail(SMP_E_INVALID, "mom_order is required");
    std::unique_ptr<smp_problem> h(new smp_problem());
    h->geo.n_sys = d->n_sys;
    for (int a = 0; a < 3; ++a) {
        h->geo.n[a] = d->n[a];
        h->geo.bc[a] = d->bc[a];
        h->geo.h[a] = d->h[a];
        if (d->row_ptr[a] && d->n_sys > 0) {
            int nnz = d->row_ptr[a][d->n_sys];
            if (nnz > 0) {
                if (!d->col[a] || !d->val[a]) return fail(SMP_E_INVALID, "CSR col/val missing");
                h->geo.M[a].rp.assign(d->row_ptr[a], d->row_ptr[a] + d->n_sys + 1);
                h->geo.M[a].ci.assign(d->col[a], d->col[a] + nnz);
                h->geo.M[a].cv.assign(d->val[a], d->val[a] + nnz);
            }
        }
    }
    std::string err = smp::build_geometry(h->geo);
    if (!err.empty()) return fail(SMP_E_INVALID, err);
    h->precision = d->precision;
    h->elem = d->precision == SMP_FP64 ? 8 : 4;
    h->f_position = d->f_position;
    h->device = d->device;
    h->mom_order.assign(d->mom_order, d->mom_order + d->n_sys);
    if (d->f_order) h->f_order.assign(d->f_order, d->f_order + d->n_sys);
    h->Q.resize(d->n_sys);

    h->host_only = d->device < 0;
    if (!h->host_only) {
        int ndev = 0;
        HIP_TRY(hipGetDeviceCount(&ndev));
        if (d->device >= ndev) return fail(SMP_E_INVALID, "no such HIP device");
        HIP_TRY(hipSetDevice(d->device));
        HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        // address of smp_probe_noalloc's fixed array on this device
        int rc = resolve_scratch(d->device);
        if (rc) return rc;
    }

    // storage geometry
    for (int e = 0; e < 3; ++e) {
        h->E[e] = 1;
        if (e < h->geo.ndim) {
            int ax = h->geo.axis[e];
            int m = 1;
            for (int j = 0; j < d->n_sys; ++j) m = std::max(m, h->geo.ext_of(j, ax));
            h->E[e] = m;
        }
    }
    h->pitch0 = h->geo.ndim > 1 ? ((h->E[0] + 15) / 16) * 16 : h->E[0];
    h->slab = (int64_t)h->pitch0 * h->E[1] * h->E[2];
    h->slab = ((h->slab + 63) / 64) * 64;
    // source offsets are 32-bit dword offsets within a component's slab
    if (h->slab * (h->elem / 4) >= ((int64_t)1 << 31))
        return fail(SMP_E_INVALID, "a component holds more than 2^31 dwords");
    if (d->tile[0] || d->tile[1] || d->tile[2]) {
        err = smp::make_tile(h->geo, std::max(1, d->tile[0]), std::max(1, d->tile[1]),
                             std::max(1, d->tile[2]), h->ts);
        if (!err.empty()) return fail(SMP_E_INVALID, err);
    } else {
        h->ts = smp::choose_tile(h->geo, h->elem);
    }
    if ((h->ts.slot_elems * (h->elem / 4) + 63) / 64 > smp::CPS_MAX)
        return fail(SMP_E_INVALID, "a tile with its halo exceeds 4 KB per component");
    h->n_tiles = 1;
    for (int e = 0; e < 3; ++e) {
        h->ntile[e] = (h->E[e] + h->ts.tile[e] - 1) / h->ts.tile[e];
        h->n_tiles *= h->ntile[e];
    }
    // schedules (re-made at the first step once the staged coefficient
    // fields are known)
    int want_blocks = 1;
    if (const char *env = std::getenv("SMP_BLOCKS_PER_CU")) want_blocks = std::atoi(env);
    if (d->blocks_per_cu) want_blocks = d->blocks_per_cu;
    int cw = 8;
    if (const char *env = std::getenv("SMP_CELL_WAVES")) cw = std::atoi(env);
    if (d->cell_waves) cw = d->cell_waves;
    if (cw != 4 && cw != 8) return fail(SMP_E_INVALID, "cell_waves must be 4 or 8");
    if (h->ts.pair && h->precision != SMP_FP32)
        return fail(SMP_E_INVALID, "tiles of more than 256 cells (two cells per thread) need the "
                                   "FP32 state");
    h->cell_waves = cw;
    if (want_blocks < 1 || want_blocks > 8) return fail(SMP_E_INVALID, "blocks_per_cu must be 1..8");
    if (const char *env = std::getenv("SMP_CHECK_EVERY")) h->check_every = std::max(0, std::atoi(env));
    h->blocks_per_cu = want_blocks;
    for (int kind = 0; kind < 2; ++kind) {
        int rc = make_plan(h.get(), kind, 0);
        if (rc) return rc;
    }
    if (h->host_only) {
        *out = h.release();
        return SMP_OK;
    }
    // state
    int rc = h->dalloc(&h->U, (size_t)h->slab * h->elem * d->n_sys);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(h->U, 0, (size_t)h->slab * h->elem * d->n_sys, h->stream));
    rc = h->dalloc(&h->red_buf, sizeof(double) * (3 * 1024 + 4));
    if (rc) return rc;
    rc = h->dalloc(&h->sink, sizeof(double) * 1024);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    *out = h.release();
    return SMP_OK;
}

int smp_destroy(smp_problem *h) {
    if (!h) return SMP_OK;
    if (!h->host_only) {
        (void)hipSetDevice(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
    }
    delete h;
    return SMP_OK;
}

int smp_num_components(const smp_problem *h, int32_t *n_sys) {
    if (!h || !n_sys) return fail(SMP_E_INVALID, "null argument");
    *n_sys = h->geo.n_sys;
    return SMP_OK;
}

int smp_grid_types(const smp_problem *h, int32_t *gt) {
    if (!h || !gt) return fail(SMP_E_INVALID, "null argument");
    std::memcpy(gt, h->geo.gt.data(), sizeof(int32_t) * 3 * h->geo.n_sys);
    return SMP_OK;
}

int smp_component_shape(const smp_problem *h, int32_t j, int64_t shape[3]) {
    if (!h || !shape) return fail(SMP_E_INVALID, "null argument");
    if (j < 0 || j >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    for (int d = 0; d < 3; ++d) shape[d] = h->geo.ext_of(j, d);
    return SMP_OK;
}

int smp_set_state(smp_problem *h, int32_t j, const double *src, int64_t count) {
    if (!h || !src) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    if (j < 0 || j >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    if (count != h->numel(j)) return fail(SMP_E_INVALID, "count does not match the component shape");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int gt[3] = {h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2]};
    char *dst = (char *)h->U + (size_t)j * h->slab * h->elem;
    if (h->precision == SMP_FP64) {
        std::vector<double> tmp(h->slab);
        h->pack<double>(gt, src, tmp.data());
        HIP_TRY(hipMemcpy(dst, tmp.data(), (size_t)h->slab * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(h->slab);
        h->pack<float>(gt, src, tmp.data());
        HIP_TRY(hipMemcpy(dst, tmp.data(), (size_t)h->slab * 4, hipMemcpyHostToDevice));
    }
    return SMP_OK;
}

int smp_get_state(smp_problem *h, int32_t j, double *dst, int64_t count) {
    if (!h || !dst) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    if (j < 0 || j >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    if (count != h->numel(j)) return fail(SMP_E_INVALID, "count does not match the component shape");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int gt[3] = {h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2]};
    const char *src = (const char *)h->U + (size_t)j * h->slab * h->elem;
    if (h->precision == SMP_FP64) {
        std::vector<double> tmp(h->slab);
        HIP_TRY(hipMemcpy(tmp.data(), src, (size_t)h->slab * 8, hipMemcpyDeviceToHost));
        h->unpack<double>(gt, tmp.data(), dst);
    } else {
        std::vector<float> tmp(h->slab);
        HIP_TRY(hipMemcpy(tmp.data(), src, (size_t)h->slab * 4, hipMemcpyDeviceToHost));
        h->unpack<float>(gt, tmp.data(), dst);
    }
    return SMP_OK;
}

int smp_set_coeff(smp_problem *h, int32_t kind, const int32_t gt_in[3], int32_t order,
                  int32_t slot, const double *field, double scalar, int64_t count) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    int gt[3] = {1, 1, 1};
    REQUIRE_DEVICE(h);
    if (kind == SMP_COEF_SIGMA_T || kind == SMP_COEF_FILTER) {
        if (!gt_in) return fail(SMP_E_INVALID, "grid type required");
        for (int d = 0; d < 3; ++d) {
            if (gt_in[d] != 1 && gt_in[d] != 2) return fail(SMP_E_INVALID, "grid type must be 1 or 2");
            gt[d] = gt_in[d];
        }
    }
    Field *f = nullptr;
    switch (kind) {
        case SMP_COEF_SIGMA_A1:
            f = &h->sA1;
            h->decay_dirty = true;
            break;
        case SMP_COEF_SIGMA_T:
            f = &h->sT[Key(gt[0], gt[1], gt[2], order, 0)];
            h->decay_dirty = true;
            break;
        case SMP_COEF_SOURCE:
            if (order < 0 || order >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
            f = &h->Q[order];
            for (int d = 0; d < 3; ++d) gt[d] = h->geo.gt[3 * order + d];
            if (!field) {
                // Q = 0 (scalar sources are expanded to a field)
                if (scalar == 0.0) {
                    f->set = false;
                    h->coef_dirty = true;
                    return SMP_OK;
                }
            }
            break;
        case SMP_COEF_FILTER:
            if (slot < 0 || slot > 2) return fail(SMP_E_INVALID, "filter slot must be 0, 1 or 2");
            f = &h->F[Key(gt[0], gt[1], gt[2], order, slot)];
            break;
        default:
            return fail(SMP_E_INVALID, "unknown coefficient kind");
    }
    const int64_t need = h->numel_gt(gt);
    if (field && count != need) return fail(SMP_E_INVALID, "field count does not match its grid");
    if (field) {
        int rc = h->upload_field(gt, field, &f->dev);
        if (rc) return rc;
    } else if (kind == SMP_COEF_SOURCE) {
        std::vector<double> tmp(need, scalar);
        int rc = h->upload_field(gt, tmp.data(), &f->dev);
        if (rc) return rc;
    } else {
        if (f->dev) {
            h->dfree(f->dev);
            f->dev = nullptr;
        }
        f->scalar = scalar;
    }
    f->set = true;
    h->coef_dirty = true;
    return SMP_OK;
}

}  // extern "C"

namespace {

template <typename T>
int refresh_decay(smp_problem *h, double dt) {
    auto one = [&](const Field &sig, Field &E, const int gt[3]) -> int {
        (void)gt;
        if (sig.dev) {
            if (!E.dev) {
                int rc = h->dalloc(&E.dev, (size_t)h->slab * sizeof(T));
                if (rc) return rc;
            }
            HIP_TRY(smp::launch_decay<T>((const T *)sig.dev, (T *)E.dev, h->slab, dt, h->stream));
        } else {
            if (E.dev) {
                h->dfree(E.dev);
                E.dev = nullptr;
            }
            E.scalar = expm1div_host(-sig.scalar * dt / 2);
        }
        E.set = true;
        return SMP_OK;
    };
    if (!h->sA1.set) return fail(SMP_E_STATE, "sigma_a1 not set");
    const int g111[3] = {1, 1, 1};
    int rc = one(h->sA1, h->EA, g111);
    if (rc) return rc;
    for (auto &kv : h->sT) {
        const int gt[3] = {std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)};
        rc = one(kv.second, h->ET[kv.first], gt);
        if (rc) return rc;
    }
    h->dt_decay = dt;
    h->decay_dirty = false;
    h->coef_dirty = true;
    return SMP_OK;
}

template <typename T>
int refresh_coefs(smp_problem *h) {
    const int n = h->geo.n_sys;
    struct CF {
        const Field *sig, *E, *Q;
        int filt;
        const void *F[2];
        double F_s[2];
    };
    std::vector<CF> cf(n);
    for (int j = 0; j < n; ++j) {
        CF &c = cf[j];
        if (j == 0) {
            c.sig = &h->sA1;
            c.E = &h->EA;
        } else {
            Key k(h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2], h->mom_order[j], 0);
            auto it = h->sT.find(k);
            if (it == h->sT.end() || !it->second.set) {
                return fail(SMP_E_STATE, "sigma_t not set for the grid type and order of component " +
                                             std::to_string(j));
            }
            c.sig = &it->second;
            c.E = &h->ET[k];
        }
        c.Q = h->Q[j].set ? &h->Q[j] : nullptr;
        c.F[0] = c.F[1] = nullptr;
        c.F_s[0] = c.F_s[1] = 1.0;
        c.filt = 0;
        if (j > 0 && h->f_position == SMP_FILTER_SUBSTEP) {
            c.filt = 1;
            for (int s = 0; s < 2; ++s) {
                Key k(h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2], h->f_order[j], s + 1);
                auto it = h->F.find(k);
                if (it == h->F.end() || !it->second.set)
                    return fail(SMP_E_STATE, "substep filter not set for component " + std::to_string(j));
                c.F[s] = it->second.dev;
                c.F_s[s] = it->second.scalar;
            }
        } else if (j > 0 && h->f_position == SMP_FILTER_FULL) {
            Key k(h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2], h->f_order[j], 0);
            auto it = h->F.find(k);
            if (it == h->F.end() || !it->second.set)
                return fail(SMP_E_STATE, "full filter not set for component " + std::to_string(j));
        }
    }
    for (int kind = 0; kind < 2; ++kind) {
        // re-plan when the number of coefficient fields to stage per tile
        // differs from the LDS the schedule reserved for them
        std::set<const void *> distinct;
        for (const smp::TgtRec &r : h->plan[kind].targets) {
            const CF &c = cf[r.comp];
            for (const Field *f : {c.sig, c.E, c.Q})
                if (f && f->dev) distinct.insert(f->dev);
        }
        const int need = std::min<int>((int)distinct.size(), smp::MAX_CFIELDS);
        if (need != h->ncf_cap[kind]) {
            int rc = make_plan(h, kind, need);
            if (rc) return rc;
        }
        // coefficient fields of this kind's targets, staged per tile in LDS
        // up to the reserved count, read from global memory beyond it
        std::vector<const void *> table;
        auto place = [&](const Field *f, int32_t &off, const void *&p, double &sc) {
            p = nullptr;
            sc = 0.0;
            if (!f) {
                off = -1;
                return;
            }
            if (!f->dev) {
                off = -1;
                sc = f->scalar;
                return;
            }
            for (size_t c = 0; c < table.size(); ++c)
                if (table[c] == f->dev) {
                    off = (int32_t)(c * h->ts.threads);
                    return;
                }
            if ((int)table.size() < h->ncf_cap[kind]) {
                off = (int32_t)(table.size() * h->ts.threads);
                table.push_back(f->dev);
                return;
            }
            off = -2;
            p = f->dev;
        };
        std::vector<smp::TgtRec> recs = h->plan[kind].targets;
        for (smp::TgtRec &r : recs) {
            const CF &c = cf[r.comp];
            double qs = 0.0;
            place(c.sig, r.sig_off, r.sig_p, r.sig_s);
            place(c.E, r.E_off, r.E_p, r.E_s);
            place(c.Q, r.Q_off, r.Q_p, qs);
            r.filt = c.filt;
            r.F[0] = c.F[0];
            r.F[1] = c.F[1];
            r.F_s[0] = c.F_s[0];
            r.F_s[1] = c.F_s[1];
        }
        h->ncf[kind] = (int)table.size();
        // device form of the records: byte offsets from the start of the
        // kernel's LDS (blob ring, staged coefficient fields, source slots)
        // and of each component in the state
        const int32_t slots_base =
            (int32_t)(smp::substep_lds_overhead() + (size_t)h->ncf[kind] * h->ts.threads * h->elem);
        for (smp::TgtRec &r : recs) {
            r.uoff = (int64_t)r.comp * h->slab * h->elem;
            for (int q = 0; q < smp::MAXT; ++q) {
                r.t[q].off = slots_base + r.t[q].off * h->elem;
                r.t[q].stride *= h->elem;
                if (h->elem == 4) {
                    const float c = (float)r.t[q].coef;
                    r.t[q].coef32[0] = c;
                    r.t[q].coef32[1] = 0.0f;
                }
            }
        }
        h->gen[kind] = false;
        for (smp::TgtRec &r : recs) {
            const bool glob = r.sig_off == -2 || r.E_off == -2 || r.Q_off == -2 || r.F[0] || r.F[1];
            if (glob) {
                h->gen[kind] = true;
            } else if (h->elem == 4) {
                // the scalar factors of the FP32 half-step: dt/2 sigma E,
                // dt/2 E, 1-F of the filter slots, sigma
                const double hd = h->dt_decay / 2;
                r.f32[0] = (float)(hd * r.sig_s * r.E_s);
                r.f32[1] = (float)(hd * r.E_s);
                r.f32[2] = (float)(1.0 - r.F_s[0]);
                r.f32[3] = (float)(1.0 - r.F_s[1]);
                r.f32[4] = (float)r.sig_s;
                r.f32[5] = 0.0f;
            }
        }
        if (!h->cf_dev[kind]) {
            int rc = h->dalloc(&h->cf_dev[kind], sizeof(void *) * smp::MAX_CFIELDS);
            if (rc) return rc;
        }
        if (!table.empty())
            HIP_TRY(hipMemcpy(h->cf_dev[kind], table.data(), sizeof(void *) * table.size(),
                              hipMemcpyHostToDevice));
        const smp::Plan &p = h->plan[kind];
        std::vector<unsigned char> blobs(std::max<size_t>(1, p.groups.size()) * smp::BLOB_SZ, 0);
        for (size_t q = 0; q < p.groups.size(); ++q) {
            const smp::DevGroup &gr = p.groups[q];
            unsigned char *b = blobs.data() + q * smp::BLOB_SZ;
            const int32_t hdr[4] = {gr.l1 - gr.l0, gr.t1 - gr.t0, 0, 0};
            std::memcpy(b, hdr, sizeof(hdr));
            std::memcpy(b + smp::BLOB_REC_OFF, recs.data() + gr.t0,
                        sizeof(smp::TgtRec) * (gr.t1 - gr.t0));
            // unused records: no cells (extent 0), every term in the zero tile
            for (int i = gr.t1 - gr.t0; i < smp::GMAX; ++i) {
                smp::TgtRec z;
                std::memset(&z, 0, sizeof(z));
                for (int q = 0; q < smp::MAXT; ++q) z.t[q].off = slots_base;
                std::memcpy(b + smp::BLOB_REC_OFF + i * sizeof(smp::TgtRec), &z, sizeof(z));
            }
            std::memcpy(b + smp::BLOB_LD_OFF, p.loads.data() + gr.l0,
                        sizeof(smp::DevLoad) * (gr.l1 - gr.l0));
        }
        HIP_TRY(hipMemcpy(h->plan_dev[kind], blobs.data(), blobs.size(), hipMemcpyHostToDevice));
    }
    h->coef_dirty = false;
    return SMP_OK;
}

template <typename T>
smp::SubstepArgs<T> make_args(smp_problem *h, int kind, double dt, int napp, int fs0, int fs1,
                              void *U = nullptr) {
    smp::SubstepArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.U = (T *)(U ? U : h->U);
    a.out = a.U;
    a.src = a.U;
    a.sink = (T *)h->sink;
    a.slab = h->slab;
    a.pitch0 = h->pitch0;
    a.ext1 = h->E[1];
    for (int e = 0; e < 3; ++e) {
        if (e < h->geo.ndim) {
            a.n[e] = h->geo.n[h->geo.axis[e]];
            a.bc[e] = h->geo.bc[h->geo.axis[e]];
        } else {
            a.n[e] = 1;
            a.bc[e] = 0;
        }
        a.tile[e] = h->ts.tile[e];
        a.ntile[e] = h->ntile[e];
        a.lds[e] = h->ts.lds[e];
    }
    a.slot_elems = h->ts.slot_elems;
    a.pair = h->ts.pair ? 1 : 0;
    a.lds0_shift = 0;
    while ((1 << a.lds0_shift) < h->ts.lds[0]) ++a.lds0_shift;
    a.inv_lds1 = 1.0f / (float)h->ts.lds[1];
    a.chunks = (h->ts.slot_elems * (h->elem / 4) + 63) / 64;
    a.n_tiles = h->n_tiles;
    const smp::Plan &p = h->plan[kind];
    a.blobs = (const unsigned char *)h->plan_dev[kind];
    a.ngroups = (int)p.groups.size();
    a.cfields = (const void *const *)h->cf_dev[kind];
    a.ncf = h->ncf[kind];
    for (int e = 0; e < 3; ++e) a.E[e] = h->E[e];
    a.half_dt = dt / 2;
    a.napp = napp;
    a.fslot[0] = fs0;
    a.fslot[1] = fs1;
    return a;
}

template <typename T>
int run_steps(smp_problem *h, double dt, int64_t n_steps) {
    if (h->decay_dirty || !(h->dt_decay == dt)) {
        int rc = refresh_decay<T>(h, dt);
        if (rc) return rc;
    }
    if (h->coef_dirty) {
        int rc = refresh_coefs<T>(h);
        if (rc) return rc;
    }
    // Launch sequence for n steps.  Substep 3 of a step and substep 1 of the
    // next one update the odd components twice from the same even state,
    // i.e. with the same flux W: inside one call they run as one launch with
    // two applications (filter slot k=2, then k=1), bit-identical to two
    // launches.  n steps take 2n+1 launches.
    const smp::SubstepArgs<T> odd_first = make_args<T>(h, 0, dt, 1, 0, 0);
    const smp::SubstepArgs<T> even = make_args<T>(h, 1, dt, 2, 0, 1);
    const smp::SubstepArgs<T> odd_fused = make_args<T>(h, 0, dt, 2, 1, 0);
    const smp::SubstepArgs<T> odd_last = make_args<T>(h, 0, dt, 1, 1, 1);
    const bool full = h->f_position == SMP_FILTER_FULL;
    std::vector<const smp::SubstepArgs<T> *> seq;
    seq.reserve(2 * n_steps + 1);
    for (int64_t s = 0; s < n_steps; ++s) {
        if (s == 0 || full)
            seq.push_back(&odd_first);
        else
            seq.push_back(&odd_fused);
        seq.push_back(&even);
        if (full || s == n_steps - 1) seq.push_back(&odd_last);
        if (full) seq.push_back(nullptr);   // full filter after the step
    }
    int64_t n_sub = 0;
    for (auto *p : seq) n_sub += p != nullptr;
    // Substep kernel time.  A sequence of substep launches only runs them
    // back to back, and its elapsed time is theirs.  Where the full filter's
    // scale kernels come between them, a start/end event pair around every
    // substep launch (up to a cap) separates the substep kernels; each pair
    // leaves the GPU idle for about 10 us between launches, so the pairs are
    // recorded only there.
    const bool per_launch = full;
    const int64_t timed = per_launch ? std::min<int64_t>(n_sub, 3 * 4096) : n_sub;
    while ((int64_t)h->events.size() < (per_launch ? 2 * timed + 2 : 2)) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        h->events.push_back(e);
    }
    hipEvent_t ev_start = h->events[0], ev_end = h->events[1];
    HIP_TRY(hipEventRecord(ev_start, h->stream));
    int64_t li = 0;
    for (auto *p : seq) {
        if (!p) {
            for (int j = 1; j < h->geo.n_sys; ++j) {
                Key k(h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2], h->f_order[j], 0);
                const Field &f = h->F[k];
                HIP_TRY(smp::launch_scale<T>((T *)h->U + (size_t)j * h->slab, (const T *)f.dev, f.scalar,
                                             h->slab, h->stream));
            }
            continue;
        }
        const int kind = p == &even ? 1 : 0;
        if (per_launch && li < timed) HIP_TRY(hipEventRecord(h->events[2 + 2 * li], h->stream));
        HIP_TRY(smp::launch_substep<T>(*p, h->plan[kind].max_terms, h->gen[kind], h->cell_waves,
                                       h->lds_bytes[kind], h->stream));
        if (per_launch && li < timed) HIP_TRY(hipEventRecord(h->events[3 + 2 * li], h->stream));
        ++li;
    }
    HIP_TRY(hipEventRecord(ev_end, h->stream));
    HIP_TRY(hipEventSynchronize(ev_end));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_end));
    h->last_total_ms = ms;
    double sub_ms = per_launch ? 0.0 : (double)ms;
    for (int64_t i = 0; per_launch && i < timed; ++i) {
        float m = 0;
        HIP_TRY(hipEventElapsedTime(&m, h->events[2 + 2 * i], h->events[3 + 2 * i]));
        sub_ms += m;
    }
    h->last_substep_ms = sub_ms;
    h->last_launches = timed;
    h->last_bytes = 0;
    // algorithmic bytes of the timed launches: every staged source component
    // read once, every target read and written once
    {
        const double elem = (double)h->elem;
        const double cells = (double)h->geo.n[0] * h->geo.n[1] * h->geo.n[2];
        const double ne = (double)h->geo.even.size(), no = (double)h->geo.odd.size();
        int64_t c = 0;
        for (auto *p : seq) {
            if (!p) continue;
            if (c++ >= timed) break;
            h->last_bytes += p == &even ? elem * cells * (no + 2 * ne) : elem * cells * (ne + 2 * no);
        }
    }
    return SMP_OK;
}

// ---------------------------------------------------------------------------
// Checkpoints of the state in device memory (slots of a full state copy)
int checkpoint_buffer(smp_problem *h, int slot, void **out) {
    if (slot < 0 || slot > 4096) return fail(SMP_E_INVALID, "checkpoint slot must be 0..4096");
    if ((int)h->ckpt.size() <= slot) h->ckpt.resize(slot + 1, nullptr);
    if (!h->ckpt[slot]) {
        int rc = h->dalloc(&h->ckpt[slot], (size_t)h->slab * h->elem * h->geo.n_sys);
        if (rc) return rc;
    }
    *out = h->ckpt[slot];
    return SMP_OK;
}

// algorithmic bytes of one substep launch of kind `kind` (0: odd targets):
// every source component read once, every target read and written once
double substep_bytes(const smp_problem *h, int kind) {
    const double cells = (double)h->geo.n[0] * h->geo.n[1] * h->geo.n[2];
    const double ne = (double)h->geo.even.size(), no = (double)h->geo.odd.size();
    return (double)h->elem * cells * (kind == 1 ? no + 2 * ne : ne + 2 * no);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of n time steps

// The LDS-staged adjoint update (adj_lds_kernel), with SMP_ADJ_KERNEL=lds:
// plans of both parities on the kernel's compiled-in tile, device records
// with the coefficients of this dt, the coefficient fields staged per tile
// (an error if the problem does not fit).  By default the gather kernels
// run: at the 2D P15 512^2 benchmark they are faster (0.28 against 0.43 ms
// per update launch), the LDS kernel being bound by the latency of its long
// per-target chains at two or three cell waves per SIMD.
template <typename T>
int adjoint_lds_setup(smp_problem *h, const std::vector<smp::AdjComp<T>> &comp) {
    h->adj_lds_ok = false;
    const char *env = std::getenv("SMP_ADJ_KERNEL");
    if (!env || std::string(env) != "lds") return SMP_OK;
    const bool strict = true;
    auto unavailable = [&](const std::string &why) {
        return strict ? fail(SMP_E_INVALID, "LDS adjoint update unavailable: " + why) : SMP_OK;
    };
    const smp::Geometry &g = h->geo;
    // the tile shape the kernel is compiled for
    smp::TileShape ts;
    if (!smp::make_tile(g, smp::adj_tile_dim(g.ndim, 0), smp::adj_tile_dim(g.ndim, 1),
                        smp::adj_tile_dim(g.ndim, 2), ts).empty() || ts.pair)
        return unavailable("tile shape");
    int bpc = 1;
    if (const char *b = std::getenv("SMP_ADJ_BLOCKS_PER_CU")) bpc = std::max(1, std::atoi(b));
    const size_t slot_bytes = (size_t)ts.slot_elems * h->elem;
    for (int kind = 0; kind < 2; ++kind) {
        const std::vector<int32_t> &cs = kind == 0 ? g.odd : g.even;
        // coefficient fields of the parity's targets
        std::vector<const void *> table;
        for (int j : cs)
            for (const void *p : {(const void *)comp[j].sig, (const void *)comp[j].E,
                                  (const void *)comp[j].dE, (const void *)comp[j].Q})
                if (p && std::find(table.begin(), table.end(), p) == table.end()) table.push_back(p);
        // the first MAX_CFIELDS are staged per tile, the rest read from
        // global memory
        if ((int)table.size() > smp::MAX_CFIELDS + 32) return unavailable("too many coefficient fields");
        std::vector<const void *> gtable;
        if ((int)table.size() > smp::MAX_CFIELDS) {
            gtable.assign(table.begin() + smp::MAX_CFIELDS, table.end());
            table.resize(smp::MAX_CFIELDS);
        }
        std::set<int> codes;
        for (int j : cs) codes.insert(g.gt[3 * j] * 9 + g.gt[3 * j + 1] * 3 + g.gt[3 * j + 2]);
        const int ngt_of = (int)codes.size();
        if (ngt_of > 4) return unavailable("more than 4 grid types of a parity");
        smp::AdjPlan plan;
        std::string err = "no LDS budget";
        for (int blocks = bpc; blocks >= 1; --blocks) {
            const size_t per_block = (160 * 1024 - 256) / blocks;
            const size_t fixed = smp::adj_lds_overhead((int)table.size(), ngt_of, h->elem);
            if (per_block <= fixed + 2 * slot_bytes) continue;
            err = smp::build_adj_plan(g, ts, kind, (int)((per_block - fixed) / slot_bytes), h->adj_recw,
                                      plan);
            if (err.empty()) break;
        }
        if (!err.empty()) return unavailable(err);
        if (plan.gcode.size() > 4) return unavailable("more than 4 grid types of a parity");
        const int ngt = (int)plan.gcode.size();
        const int ncf = (int)table.size();
        const size_t base = smp::adj_lds_overhead(ncf, ngt, h->elem);
        const int32_t slots_base = (int32_t)base;
        auto field_off = [&](const void *p) -> int32_t {
            if (!p) return -1;
            const auto it = std::find(table.begin(), table.end(), p);
            if (it != table.end()) return (int32_t)((it - table.begin()) * 256);
            return (int32_t)(-2 - (std::find(gtable.begin(), gtable.end(), p) - gtable.begin()));
        };
        // device records (2 * ADJ_TPH per group; positions without a target
        // have no points, every term in the zero tile of slot 0): offsets in
        // bytes from the start of LDS, the coefficients of this dt
        std::vector<smp::AdjRec> recs = plan.targets;
        if (recs.empty()) recs.resize(1);
        std::vector<unsigned char> blobs(std::max<size_t>(1, plan.groups.size()) * smp::BLOB_SZ, 0);
        for (size_t q = 0; q < plan.groups.size(); ++q) {
            const smp::DevGroup &gr = plan.groups[q];
            unsigned char *bl = blobs.data() + q * smp::BLOB_SZ;
            const int32_t hdr[4] = {gr.l1 - gr.l0, gr.t1 - gr.t0, 0, 0};
            std::memcpy(bl, hdr, sizeof(hdr));
            std::memcpy(bl + smp::ADJ_BLOB_LD_OFF, plan.loads.data() + gr.l0,
                        sizeof(smp::DevLoad) * (gr.l1 - gr.l0));
        }
        for (smp::AdjRec &r : recs) {
            for (int k = 0; k < 2 * smp::MAXT; ++k) r.off[k] = slots_base + r.off[k] * h->elem;
            if (r.comp < 0) {
                r.sig_off = r.E_off = r.dE_off = r.Q_off = -1;
                continue;
            }
            {
                const smp::AdjComp<T> &c = comp[r.comp];
                r.sig_off = field_off(c.sig);
                r.E_off = field_off(c.E);
                r.dE_off = field_off(c.dE);
                r.Q_off = field_off(c.Q);
                r.sig_s = c.sig_s;
                r.E_s = c.E_s;
                r.dE_s = c.dE_s;
                r.F_s[0] = c.F[0];
                r.F_s[1] = c.F[1];
                r.uoff = (int64_t)r.comp * h->slab * h->elem;
            }
        }
        for (void **p : {&h->adj_lds_recs[kind], &h->adj_lds_blobs[kind], &h->adj_lds_cf[kind],
                         &h->adj_lds_gf[kind]}) {
            h->dfree(*p);
            *p = nullptr;
        }
        int rc = h->dalloc(&h->adj_lds_recs[kind], recs.size() * sizeof(smp::AdjRec));
        if (!rc) rc = h->dalloc(&h->adj_lds_blobs[kind], blobs.size());
        if (!rc) rc = h->dalloc(&h->adj_lds_cf[kind], sizeof(void *) * smp::MAX_CFIELDS);
        if (!rc) rc = h->dalloc(&h->adj_lds_gf[kind], sizeof(void *) * std::max<size_t>(1, gtable.size()));
        if (rc) return rc;
        if (!gtable.empty())
            HIP_TRY(hipMemcpy(h->adj_lds_gf[kind], gtable.data(), sizeof(void *) * gtable.size(),
                              hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(h->adj_lds_recs[kind], recs.data(), recs.size() * sizeof(smp::AdjRec),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(h->adj_lds_blobs[kind], blobs.data(), blobs.size(), hipMemcpyHostToDevice));
        if (!table.empty())
            HIP_TRY(hipMemcpy(h->adj_lds_cf[kind], table.data(), sizeof(void *) * table.size(),
                              hipMemcpyHostToDevice));
        h->adj_plan[kind] = plan;
        h->adj_lds_ncf[kind] = ncf;
        h->adj_lds_gen[kind] = !gtable.empty();
        h->adj_lds_bytes[kind] = base + (size_t)std::max(plan.n_slots, 1) * slot_bytes;
    }
    if (!h->adj_zeros) {
        int rc = h->dalloc(&h->adj_zeros, 256);
        if (rc) return rc;
        HIP_TRY(hipMemset(h->adj_zeros, 0, 256));
    }
    h->adj_ts = ts;
    h->adj_ntiles = 1;
    for (int e = 0; e < 3; ++e) {
        h->adj_ntile[e] = (h->E[e] + ts.tile[e] - 1) / ts.tile[e];
        h->adj_ntiles *= h->adj_ntile[e];
    }
    h->adj_lds_ok = true;
    return SMP_OK;
}

// device tables of the adjoint kernels (coefficients as the forward uses
// them at the current dt)
template <typename T>
int adjoint_tables(smp_problem *h, double dt) {
    const smp::Geometry &g = h->geo;
    std::set<const void *> fresh;   // dE/dsigma fields computed for this dt
    const int n = g.n_sys;
    int kaxis_of[3] = {-1, -1, -1};
    for (int e = 0; e < g.ndim; ++e) kaxis_of[g.axis[e]] = e;
    std::vector<smp::AdjComp<T>> comp(n);
    std::vector<smp::AdjTerm> rows, cols;
    for (int j = 0; j < n; ++j) {
        smp::AdjComp<T> &c = comp[j];
        std::memset(&c, 0, sizeof(c));
        const Field *sig, *E;
        if (j == 0) {
            sig = &h->sA1;
            E = &h->EA;
        } else {
            Key k(g.gt[3 * j], g.gt[3 * j + 1], g.gt[3 * j + 2], h->mom_order[j], 0);
            auto it = h->sT.find(k);
            if (it == h->sT.end() || !it->second.set)
                return fail(SMP_E_STATE, "sigma_t not set for component " + std::to_string(j));
            sig = &it->second;
            E = &h->ET[k];
        }
        c.sig = (const T *)sig->dev;
        c.sig_s = sig->scalar;
        c.E = (const T *)E->dev;
        c.E_s = E->scalar;
        if (sig->dev) {
            void *&d = h->adj_dE[sig->dev];
            if (!d) {
                int rc = h->dalloc(&d, (size_t)h->slab * sizeof(T));
                if (rc) return rc;
            }
            if (fresh.insert(sig->dev).second)
                HIP_TRY(smp::launch_dedsig<T>((const T *)sig->dev, (T *)d, h->slab, dt, h->stream));
            c.dE = (const T *)d;
        } else {
            const double x = -sig->scalar * dt / 2;
            c.dE_s = -dt / 2 * (std::fabs(x) > 2e-4 ? (std::exp(x) * (x - 1) + 1) / (x * x) : .5 + x / 3);
        }
        c.Q = h->Q[j].set ? (const T *)h->Q[j].dev : nullptr;
        if (h->Q[j].set && !h->Q[j].dev && h->Q[j].scalar != 0.0)
            return fail(SMP_E_INVALID, "the adjoint needs sources as fields");
        c.F[0] = c.F[1] = 1.0;
        if (j > 0 && h->f_position == SMP_FILTER_SUBSTEP) {
            for (int sl = 0; sl < 2; ++sl) {
                Key k(g.gt[3 * j], g.gt[3 * j + 1], g.gt[3 * j + 2], h->f_order[j], sl + 1);
                auto it = h->F.find(k);
                if (it == h->F.end() || !it->second.set)
                    return fail(SMP_E_STATE, "substep filter not set for component " + std::to_string(j));
                if (it->second.dev)
                    return fail(SMP_E_INVALID, "the adjoint supports filters that are constant in space");
                c.F[sl] = it->second.scalar;
            }
        }
        for (int e = 0; e < 3; ++e) {
            c.ext[e] = e < g.ndim ? g.ext_of(j, g.axis[e]) : 1;
            c.cell[e] = e < g.ndim && g.gt[3 * j + g.axis[e]] == 1;
        }
    }
    // terms of each component by kernel axis, each active axis padded to
    // whole groups (at least one) with coefficient 0 terms that sample the
    // axis list's first component, or the list's own component when the
    // axis list is empty
    std::vector<std::vector<std::vector<smp::AdjTerm>>> rows_by(n), cols_by(n);
    for (int j = 0; j < n; ++j) {
        rows_by[j].resize(g.ndim);
        cols_by[j].resize(g.ndim);
    }
    for (int j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) {
            const smp::Csr &M = g.M[d];
            if (M.rp.empty()) continue;
            const int e = kaxis_of[d];
            for (int q = M.rp[j]; q < M.rp[j + 1]; ++q) {
                const double coef = -M.cv[q] / g.h[d];
                smp::AdjTerm t;
                std::memset(&t, 0, sizeof(t));
                t.k = M.ci[q];
                t.coef = coef;
                rows_by[j][e].push_back(t);
                t.k = j;
                cols_by[M.ci[q]][e].push_back(t);
            }
        }
    auto flatten = [&](std::vector<std::vector<smp::AdjTerm>> &by, std::vector<smp::AdjTerm> &out,
                       int own, int32_t *ra) {
        for (int e = 0; e < 4; ++e) ra[e] = (int32_t)out.size();
        for (int e = 0; e < g.ndim; ++e) {
            ra[e] = (int32_t)out.size();
            std::vector<smp::AdjTerm> &v = by[e];
            smp::AdjTerm pad;
            std::memset(&pad, 0, sizeof(pad));
            pad.k = v.empty() ? own : v[0].k;
            while (v.empty() || v.size() % smp::ADJ_GROUP) v.push_back(pad);
            out.insert(out.end(), v.begin(), v.end());
        }
        for (int e = g.ndim; e < 4; ++e) ra[e] = (int32_t)out.size();
    };
    for (int j = 0; j < n; ++j) {
        flatten(rows_by[j], rows, j, comp[j].ra);
        flatten(cols_by[j], cols, j, comp[j].ca);
    }
    // components of each parity grouped by grid type: odd first, then even
    std::vector<int32_t> list;
    for (int par = 0; par < 2; ++par) {
        const std::vector<int32_t> &cs = par == 0 ? g.odd : g.even;
        std::vector<int> codes;
        for (int j : cs) {
            const int code = (g.gt[3 * j] - 1) * 4 + (g.gt[3 * j + 1] - 1) * 2 + (g.gt[3 * j + 2] - 1);
            if (std::find(codes.begin(), codes.end(), code) == codes.end()) codes.push_back(code);
        }
        h->adj_ngrp[par] = (int)codes.size();
        const int base = (int)list.size();
        for (size_t q = 0; q < codes.size(); ++q) {
            h->adj_grp0[par][q] = (int32_t)list.size() - base;
            h->adj_gcode[par][q] = codes[q];
            for (int j : cs) {
                const int code = (g.gt[3 * j] - 1) * 4 + (g.gt[3 * j + 1] - 1) * 2 + (g.gt[3 * j + 2] - 1);
                if (code == codes[q]) list.push_back(j);
            }
        }
        h->adj_grp0[par][codes.size()] = (int32_t)list.size() - base;
    }
    for (void **p : {&h->adj_comp, &h->adj_rows, &h->adj_cols, &h->adj_list}) {
        h->dfree(*p);
        *p = nullptr;
    }
    int rc = h->dalloc(&h->adj_comp, sizeof(smp::AdjComp<T>) * n);
    if (!rc) rc = h->dalloc(&h->adj_rows, sizeof(smp::AdjTerm) * std::max<size_t>(1, rows.size()));
    if (!rc) rc = h->dalloc(&h->adj_cols, sizeof(smp::AdjTerm) * std::max<size_t>(1, cols.size()));
    if (!rc) rc = h->dalloc(&h->adj_list, sizeof(int32_t) * std::max<size_t>(1, list.size()));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(h->adj_comp, comp.data(), sizeof(smp::AdjComp<T>) * n, hipMemcpyHostToDevice));
    if (!rows.empty())
        HIP_TRY(hipMemcpy(h->adj_rows, rows.data(), sizeof(smp::AdjTerm) * rows.size(),
                          hipMemcpyHostToDevice));
    if (!cols.empty())
        HIP_TRY(hipMemcpy(h->adj_cols, cols.data(), sizeof(smp::AdjTerm) * cols.size(),
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->adj_list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice));
    return adjoint_lds_setup<T>(h, comp);
}

int adjoint_buffers(smp_problem *h) {
    const size_t st = (size_t)h->slab * h->elem * h->geo.n_sys;
    for (void **p : {&h->lam, &h->mu}) {
        if (!*p) {
            int rc = h->dalloc(p, st);
            if (rc) return rc;
            HIP_TRY(hipMemsetAsync(*p, 0, st, h->stream));
        }
    }
    if (!h->G) {
        // dJ/dsigma_a and dJ/dsigma_s0 for each of the 8 grid type codes
        int rc = h->dalloc(&h->G, (size_t)h->slab * 8 * 16);
        if (rc) return rc;
        HIP_TRY(hipMemsetAsync(h->G, 0, (size_t)h->slab * 8 * 16, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

template <typename T>
int run_adjoint(smp_problem *h, double dt, int64_t n_steps, int every) {
    if (h->f_position == SMP_FILTER_FULL)
        return fail(SMP_E_INVALID, "the adjoint supports no filter or the substep filter");
    // the adjoint kernels address a component with 32-bit byte offsets
    if ((double)h->slab * h->elem >= 2147483648.0)
        return fail(SMP_E_INVALID, "the adjoint supports components of less than 2 GiB");
    int rc = adjoint_buffers(h);
    if (rc) return rc;
    // the forward coefficients at this dt
    if (h->decay_dirty || !(h->dt_decay == dt)) {
        rc = refresh_decay<T>(h, dt);
        if (rc) return rc;
    }
    if (h->coef_dirty) {
        rc = refresh_coefs<T>(h);
        if (rc) return rc;
    }
    // Recording launches also store W+Q of their targets, so that the update
    // needs no forward flux of its own (the gather update: 16 fewer gathers
    // per target in 2D, about 25% less time per launch; the LDS update: no
    // state tiles), for twice the memory per recorded segment (under a
    // budget, fewer segments recorded in the forward pass).  Not with
    // global-memory coefficients or with SMP_ADJ_RECORD_W=0.
    const char *rw_env = std::getenv("SMP_ADJ_RECORD_W");
    const bool recW = !h->gen[0] && !h->gen[1] && !(rw_env && std::string(rw_env) == "0");
    h->adj_recw = recW;
    rc = adjoint_tables<T>(h, dt);
    if (rc) return rc;
    const size_t st = (size_t)h->slab * h->elem * h->geo.n_sys;
    smp::AdjSchedule sc;
    {
        const int64_t nbuf = h->adj_budget > 0 ? h->adj_budget / (int64_t)st : -1;
        const std::string err = smp::plan_adjoint(n_steps, every, nbuf, recW ? 2 : 1, sc);
        if (!err.empty()) return fail(h->adj_budget > 0 ? SMP_E_OOM : SMP_E_INVALID, err);
    }
    const int64_t k = sc.seg_steps, rb = sc.rec_factor * (k + 1);
    const int H = sc.segments - sc.tail;
    // buffers of the schedule, kept on the handle for the next call
    while ((int)h->adj_slot.size() < sc.slots) {
        void *p = nullptr;
        rc = h->dalloc(&p, st);
        if (rc) return rc;
        h->adj_slot.push_back(p);
    }
    if (h->adj_range_rb != rb) {
        for (void *p : h->adj_range) h->dfree(p);
        h->adj_range.clear();
        h->adj_range_rb = rb;
    }
    while ((int)h->adj_range.size() < sc.ranges) {
        void *p = nullptr;
        rc = h->dalloc(&p, st * (size_t)rb);
        if (rc) return rc;
        h->adj_range.push_back(p);
    }
    if (sc.work && !h->adj_work) {
        rc = h->dalloc(&h->adj_work, st);
        if (rc) return rc;
    }
    if (sc.tail > 0 && !h->adj_tstart) {
        rc = h->dalloc(&h->adj_tstart, st);
        if (rc) return rc;
    }
    h->adj_sched = sc;
    h->adj_peak_bytes = sc.buffers * (int64_t)st;
    for (int q = 0; q < 8; ++q)
        if (!h->adj_ev[q])
            HIP_TRY(q < 3 ? hipEventCreate(&h->adj_ev[q])
                          : hipEventCreateWithFlags(&h->adj_ev[q], hipEventDisableTiming));
    while ((int)h->adj_range_ev.size() < 2 * sc.ranges) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        h->adj_range_ev.push_back(e);
    }
    hipEvent_t e0 = h->adj_ev[0], e1 = h->adj_ev[1], e2 = h->adj_ev[2], ev_fwd = h->adj_ev[3];
    if (!h->adj_stream) HIP_TRY(hipStreamCreateWithFlags(&h->adj_stream, hipStreamNonBlocking));
    hipStream_t rs = h->adj_stream;
    const smp::SubstepArgs<T> f_odd_first = make_args<T>(h, 0, dt, 1, 0, 0);
    const smp::SubstepArgs<T> f_even = make_args<T>(h, 1, dt, 2, 0, 1);
    const smp::SubstepArgs<T> f_odd_fused = make_args<T>(h, 0, dt, 2, 1, 0);
    const smp::SubstepArgs<T> f_odd_last = make_args<T>(h, 0, dt, 1, 1, 1);
    // A segment of m steps is 2m+1 launches (odd first; the even substep and
    // the fused odd substeps of consecutive steps; odd last): launch L updates
    // the targets of one parity (odd for even L) from the other parity.
    auto seg_len = [&](int64_t sg) { return std::min<int64_t>(k, n_steps - sg * k); };
    auto nl_of = [&](int64_t sg) { return 2 * seg_len(sg) + 1; };
    auto launch_of = [&](int64_t L, int64_t nl) -> const smp::SubstepArgs<T> & {
        if (L == 0) return f_odd_first;
        if (L == nl - 1) return f_odd_last;
        return L % 2 ? f_even : f_odd_fused;
    };
    // record L of a range: the targets of launch L after it; records 2b and
    // 2b+1 (odd and even targets) share buffer b of the range
    auto rec = [&](int r, int64_t L) -> void * {
        return (char *)h->adj_range[r] + (size_t)(L / 2) * st;
    };
    // W+Q of launch L's targets, in the second half of the range
    auto wrec = [&](int r, int64_t L) -> void * {
        return recW ? (char *)h->adj_range[r] + (size_t)(k + 1 + L / 2) * st : nullptr;
    };
    auto buf_of = [&](int id) -> void * {
        if (id >= 0) return h->adj_slot[id];
        if (id == smp::ADJ_WORK) return h->adj_work;
        if (id == smp::ADJ_TAIL_START) return h->adj_tstart;
        return nullptr;
    };
    // where each parity of a segment's start state lives: a full state, or
    // the previous (recorded) segment's last two records
    struct Start {
        void *odd, *even;
    };
    auto start_of = [&](int64_t sg, int src) -> Start {
        if (src == smp::ADJ_PREV) {
            const int64_t nlp = nl_of(sg - 1);
            const int rp = (int)(sg - 1 - H);
            return Start{rec(rp, nlp - 1), rec(rp, nlp - 2)};
        }
        void *b = buf_of(src);
        return Start{b, b};
    };
    // the state after launch m of a segment, for the parity launch m targets
    // (m = -2: odd, m = -1: even, at the segment start)
    auto after = [&](const Start &s0, int r, int64_t m) -> void * {
        if (m >= 0) return rec(r, m);
        return m == -2 ? s0.odd : s0.even;
    };
    // segment sg out of place into range r: launch L reads its targets after
    // launch L-2 and its sources after launch L-1, and writes record L
    auto record_seg = [&](int64_t sg, const Start &s0, int r, hipStream_t strm) -> int {
        const int64_t nl = nl_of(sg);
        for (int64_t L = 0; L < nl; ++L) {
            const int kind = L % 2;
            smp::SubstepArgs<T> a = launch_of(L, nl);
            a.U = (T *)after(s0, r, L - 2);
            a.src = (const T *)after(s0, r, L - 1);
            a.out = (T *)rec(r, L);
            a.wout = (T *)wrec(r, L);
            HIP_TRY(smp::launch_substep<T>(a, h->plan[kind].max_terms, h->gen[kind], h->cell_waves,
                                           h->lds_bytes[kind], strm));
        }
        return SMP_OK;
    };
    // segments [s0, s1) in place on the state in buf
    auto run_inplace = [&](int64_t s0, int64_t s1, void *buf, hipStream_t strm) -> int {
        for (int64_t sg = s0; sg < s1; ++sg) {
            const int64_t nl = nl_of(sg);
            for (int64_t L = 0; L < nl; ++L) {
                const int kind = L % 2;
                smp::SubstepArgs<T> a = launch_of(L, nl);
                a.U = a.out = (T *)buf;
                a.src = (const T *)buf;
                HIP_TRY(smp::launch_substep<T>(a, h->plan[kind].max_terms, h->gen[kind], h->cell_waves,
                                               h->lds_bytes[kind], strm));
            }
        }
        return SMP_OK;
    };
    const int nl_odd = (int)h->geo.odd.size(), nl_even = (int)h->geo.even.size();
    const int32_t *list_odd = (const int32_t *)h->adj_list;
    const int32_t *list_even = list_odd + nl_odd;
    auto adj_args = [&](const void *U, const void *U0, const void *W, const int32_t *list, int nl,
                        int napp, int f0, int f1) {
        const int par = list == list_odd ? 0 : 1;
        smp::AdjArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        a.comp = (const smp::AdjComp<T> *)h->adj_comp;
        a.rows = (const smp::AdjTerm *)h->adj_rows;
        a.cols = (const smp::AdjTerm *)h->adj_cols;
        a.list = list;
        a.n_list = nl;
        a.n_grp = h->adj_ngrp[par];
        for (int q = 0; q < 9; ++q) a.grp0[q] = h->adj_grp0[par][q];
        for (int q = 0; q < 8; ++q) a.gcode[q] = h->adj_gcode[par][q];
        a.U = (const T *)U;
        a.U0 = (const T *)U0;
        a.W = (const T *)W;
        a.lam = (T *)h->lam;
        a.mu = (T *)h->mu;
        a.Ga = (double *)h->G;
        a.Gs = (double *)h->G + 8 * h->slab;
        a.slab = h->slab;
        a.pitch0 = h->pitch0;
        a.ext1 = h->E[1];
        for (int e = 0; e < 3; ++e) {
            a.E[e] = h->E[e];
            a.n[e] = e < h->geo.ndim ? h->geo.n[h->geo.axis[e]] : 1;
            a.bc[e] = e < h->geo.ndim ? h->geo.bc[h->geo.axis[e]] : 0;
        }
        a.half_dt = dt / 2;
        a.napp = napp;
        a.fslot[0] = f0;
        a.fslot[1] = f1;
        a.ndim = h->geo.ndim;
        return a;
    };
    auto lds_args = [&](const void *U, const void *U0, const void *W, int kind, int napp, int f0, int f1) {
        smp::AdjTileArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        const smp::TileShape &ts = h->adj_ts;
        const smp::AdjPlan &p = h->adj_plan[kind];
        a.U = (const T *)U;
        a.U0 = (const T *)U0;
        a.W = (const T *)W;
        a.lam = (T *)h->lam;
        a.mu = (T *)h->mu;
        a.Ga = (double *)h->G;
        a.Gs = (double *)h->G + 8 * h->slab;
        a.sink = (T *)h->sink;
        a.slab = h->slab;
        a.pitch0 = h->pitch0;
        a.ext1 = h->E[1];
        for (int e = 0; e < 3; ++e) {
            a.n[e] = e < h->geo.ndim ? h->geo.n[h->geo.axis[e]] : 1;
            a.bc[e] = e < h->geo.ndim ? h->geo.bc[h->geo.axis[e]] : 0;
            a.tile[e] = ts.tile[e];
            a.ntile[e] = h->adj_ntile[e];
            a.lds[e] = ts.lds[e];
            a.E[e] = h->E[e];
        }
        a.slot_elems = ts.slot_elems;
        while ((1 << a.lds0_shift) < ts.lds[0]) ++a.lds0_shift;
        a.inv_lds1 = 1.0f / (float)ts.lds[1];
        a.chunks = (ts.slot_elems * (h->elem / 4) + 63) / 64;
        a.n_tiles = h->adj_ntiles;
        a.ngroups = (int)p.groups.size();
        a.blobs = (const unsigned char *)h->adj_lds_blobs[kind];
        a.recs = (const smp::AdjRec *)h->adj_lds_recs[kind];
        a.zeros = (const unsigned *)h->adj_zeros;
        a.cfields = (const void *const *)h->adj_lds_cf[kind];
        a.gfields = (const void *const *)h->adj_lds_gf[kind];
        a.ncf = h->adj_lds_ncf[kind];
        a.half_dt = dt / 2;
        a.napp = napp;
        a.fslot[0] = f0;
        a.fslot[1] = f1;
        a.ngt = (int)p.gcode.size();
        for (int q = 0; q < a.ngt; ++q) {
            const int code = p.gcode[q];
            a.gcode[q] = code;
            const int gt[3] = {code / 4 + 1, (code / 2) % 2 + 1, code % 2 + 1};
            for (int e = 0; e < 3; ++e) {
                const int d = e < h->geo.ndim ? h->geo.axis[e] : -1;
                a.gext[q][e] = d < 0 ? 1 : (gt[d] == 1 ? h->geo.n[d] : h->geo.n[d] + h->geo.bc[d]);
            }
        }
        return a;
    };
    // the adjoint of segment sg, launch by launch from the last: the update
    // of L's targets (L's targets before L and its sources are the states
    // after launches L-2 and L-1), with the transposed flux of launch L+1
    // fused in; the segment's first launch ends with the transposed flux
    // into its sources as a launch of its own
    auto reverse_seg = [&](int64_t sg, const Start &s0, int r) -> int {
        const int64_t nl = nl_of(sg);
        for (int64_t L = nl - 1; L >= 0; --L) {
            const int kind = L % 2;
            const smp::SubstepArgs<T> &f = launch_of(L, nl);
            const void *U0 = after(s0, r, L - 2);
            const void *U = after(s0, r, L - 1);
            const int32_t *tl = kind ? list_even : list_odd, *sl = kind ? list_odd : list_even;
            const int ntl = kind ? nl_even : nl_odd, nsl = kind ? nl_odd : nl_even;
            if (h->adj_lds_ok)
                HIP_TRY(smp::launch_adj_lds<T>(lds_args(U, U0, wrec(r, L), kind, f.napp, f.fslot[0], f.fslot[1]),
                                               h->geo.ndim, L + 1 < nl, h->adj_lds_gen[kind],
                                               h->adj_lds_bytes[kind], h->stream));
            else
                HIP_TRY(smp::launch_adj_update<T>(adj_args(U, U0, wrec(r, L), tl, ntl, f.napp, f.fslot[0],
                                                           f.fslot[1]),
                                                  L + 1 < nl, h->stream));
            if (L == 0)
                HIP_TRY(smp::launch_adj_transpose<T>(adj_args(U, U0, nullptr, sl, nsl, f.napp, f.fslot[0],
                                                              f.fslot[1]),
                                                     h->stream));
        }
        return SMP_OK;
    };

    // forward pass (the handle's stream): checkpoints and in-place segments
    // of the head, the tail recorded out of place, the final state
    HIP_TRY(hipEventRecord(e0, h->stream));
    for (const smp::AdjAction &a : sc.fwd) {
        switch (a.kind) {
        case smp::ADJ_FWD_RUN:
            rc = run_inplace(a.s0, a.s1, h->U, h->stream);
            break;
        case smp::ADJ_FWD_CKPT:
            HIP_TRY(hipMemcpyAsync(buf_of(a.dst), h->U, st, hipMemcpyDeviceToDevice, h->stream));
            break;
        case smp::ADJ_FWD_REC:
            rc = record_seg(a.s0, start_of(a.s0, a.src), a.range, h->stream);
            break;
        case smp::ADJ_FWD_FINAL: {
            const int64_t nl = nl_of(a.s0);
            const size_t sb = (size_t)h->slab * h->elem;
            for (int par = 0; par < 2; ++par)
                for (int j : par == 0 ? h->geo.odd : h->geo.even)
                    HIP_TRY(hipMemcpyAsync((char *)h->U + j * sb,
                                           (const char *)rec(a.range, par == 0 ? nl - 1 : nl - 2) + j * sb,
                                           sb, hipMemcpyDeviceToDevice, h->stream));
            break;
        }
        default:
            return fail(SMP_E_INVALID, "adjoint schedule: unexpected forward action");
        }
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(e1, h->stream));
    HIP_TRY(hipEventRecord(ev_fwd, h->stream));
    HIP_TRY(hipStreamWaitEvent(rs, ev_fwd, 0));
    // reverse sweep (the handle's stream) with the recomputation and the
    // re-advanced segments on the second stream: a recomputation into a
    // range waits until the range's previous segment is stepped back, a
    // reversal until its range is recomputed, and a re-advance until every
    // reversal issued before it is done (it may overwrite a start state
    // that a reversal still reads)
    std::vector<char> rc_pending(sc.ranges, 0), rev_done(sc.ranges, 0);
    hipEvent_t last_rev = nullptr;
    for (const smp::AdjAction &a : sc.rev) {
        switch (a.kind) {
        case smp::ADJ_ADVANCE:
            if (last_rev) HIP_TRY(hipStreamWaitEvent(rs, last_rev, 0));
            HIP_TRY(hipMemcpyAsync(buf_of(a.dst), buf_of(a.src), st, hipMemcpyDeviceToDevice, rs));
            rc = run_inplace(a.s0, a.s1, buf_of(a.dst), rs);
            break;
        case smp::ADJ_RECOMP: {
            if (rev_done[a.range]) HIP_TRY(hipStreamWaitEvent(rs, h->adj_range_ev[2 * a.range + 1], 0));
            rc = record_seg(a.s0, start_of(a.s0, a.src), a.range, rs);
            if (rc) break;
            HIP_TRY(hipEventRecord(h->adj_range_ev[2 * a.range], rs));
            rc_pending[a.range] = 1;
            break;
        }
        case smp::ADJ_REVERSE: {
            if (rc_pending[a.range]) {
                HIP_TRY(hipStreamWaitEvent(h->stream, h->adj_range_ev[2 * a.range], 0));
                rc_pending[a.range] = 0;
            }
            rc = reverse_seg(a.s0, start_of(a.s0, a.src), a.range);
            if (rc) break;
            HIP_TRY(hipEventRecord(h->adj_range_ev[2 * a.range + 1], h->stream));
            rev_done[a.range] = 1;
            last_rev = h->adj_range_ev[2 * a.range + 1];
            break;
        }
        default:
            return fail(SMP_E_INVALID, "adjoint schedule: unexpected reverse action");
        }
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(rs));
    HIP_TRY(hipEventRecord(e2, h->stream));
    HIP_TRY(hipEventSynchronize(e2));
    float f = 0, r = 0;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    HIP_TRY(hipEventElapsedTime(&r, e1, e2));
    h->adj_fwd_ms = f;
    h->adj_rev_ms = r;
    return SMP_OK;
}

}  // namespace

struct smp_batch {
    std::vector<smp_problem *> hs;
    int device = 0;
    int precision = SMP_FP32;
    hipStream_t stream = nullptr;
    // device tables: entries of every launch, one descriptor and one counter
    // per launch
    void *entries = nullptr;
    size_t entries_cap = 0;
    void *launch = nullptr;
    int *counters = nullptr;
    int launch_cap = 0;
    // captured launch sequence and the structure it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_launches = -1, g_grid = 0, g_nt = 0, g_cw = 0;
    size_t g_lds = 0;
    bool g_gen = false;
    std::vector<int> g_napp;   // applications of each captured launch
    std::vector<hipEvent_t> events;
    double last_ms = 0, last_kernel_ms = 0, last_bytes = 0;
    int64_t last_launches = 0, last_mcu = 0;
    int grid = 0;
    size_t lds = 0;

    ~smp_batch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : {entries, launch, (void *)counters})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
int batch_run(smp_batch *b, const double *dt, const int64_t *ns, int use_graph) {
    const int n = (int)b->hs.size();
    int64_t max_n = 0;
    for (int i = 0; i < n; ++i) {
        if (!(dt[i] > 0) || !std::isfinite(dt[i]))
            return fail(SMP_E_INVALID, "dt must be positive and finite");
        if (ns[i] < 0) return fail(SMP_E_INVALID, "n_steps must be >= 0");
        max_n = std::max(max_n, ns[i]);
    }
    // coefficients of every problem for its dt (the problems' own launches
    // are finished first)
    for (int i = 0; i < n; ++i) {
        smp_problem *h = b->hs[i];
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (ns[i] == 0) continue;
        if (h->decay_dirty || !(h->dt_decay == dt[i])) {
            int rc = refresh_decay<T>(h, dt[i]);
            if (rc) return rc;
        }
        if (h->coef_dirty) {
            int rc = refresh_coefs<T>(h);
            if (rc) return rc;
        }
    }
    // the decay kernels ran on each problem's own stream: finish them before
    // the batch's stream reads the E fields
    for (int i = 0; i < n; ++i) HIP_TRY(hipStreamSynchronize(b->hs[i]->stream));
    b->last_ms = b->last_kernel_ms = b->last_bytes = 0;
    b->last_launches = b->last_mcu = 0;
    if (max_n == 0) return SMP_OK;
    // launch L: L = 0 the first odd substep of every problem; L = 2k+1 the
    // even substep of step k; L = 2k (k >= 1) the odd substep that ends step
    // k-1 and, fused, begins step k, or ends the last step
    struct Var {
        smp::SubstepArgs<T> odd_first, even, odd_fused, odd_last;
    };
    std::vector<Var> var(n);
    int nt = 0;
    bool gen = false;
    size_t lds = 0;
    for (int i = 0; i < n; ++i) {
        smp_problem *h = b->hs[i];
        if (ns[i] == 0) continue;
        var[i].odd_first = make_args<T>(h, 0, dt[i], 1, 0, 0);
        var[i].even = make_args<T>(h, 1, dt[i], 2, 0, 1);
        var[i].odd_fused = make_args<T>(h, 0, dt[i], 2, 1, 0);
        var[i].odd_last = make_args<T>(h, 0, dt[i], 1, 1, 1);
        for (int kind = 0; kind < 2; ++kind) {
            nt = std::max(nt, h->plan[kind].max_terms);
            gen = gen || h->gen[kind];
            lds = std::max(lds, h->lds_bytes[kind]);
        }
    }
    // Launch slots: substep position L of the lockstep, split by the number
    // of half-step applications (the problems ending at L take one, the
    // others two), so that each launch runs one compiled tile path.  The
    // entries of a position are independent problems.
    const int n_pos = (int)(2 * max_n + 1);
    std::vector<smp::BatchEntry<T>> ent;
    std::vector<smp::BatchLaunch<T>> desc;
    std::vector<int> lnapp;
    std::vector<size_t> first;
    double bytes = 0;
    for (int L = 0; L < n_pos; ++L) {
        const int64_t k = L / 2;
        for (int napp = 2; napp >= 1; --napp) {
            const size_t f = ent.size();
            for (int i = 0; i < n; ++i) {
                if (ns[i] == 0) continue;
                const smp::SubstepArgs<T> *a = nullptr;
                if (L == 0)
                    a = &var[i].odd_first;
                else if (L & 1)
                    a = k < ns[i] ? &var[i].even : nullptr;
                else
                    a = k < ns[i] ? &var[i].odd_fused : (k == ns[i] ? &var[i].odd_last : nullptr);
                if (!a || a->napp != napp) continue;
                smp::BatchEntry<T> e;
                std::memset(&e, 0, sizeof(e));
                e.a = *a;
                e.pad = i;
                ent.push_back(e);
                bytes += substep_bytes(b->hs[i], (L & 1) ? 1 : 0);
            }
            if (ent.size() == f) continue;
            // heaviest tiles first (most groups per tile), so that the last
            // units handed out are the cheap ones
            std::stable_sort(ent.begin() + f, ent.end(),
                             [](const smp::BatchEntry<T> &x, const smp::BatchEntry<T> &y) {
                                 return x.a.ngroups > y.a.ngroups;
                             });
            int32_t u = 0;
            for (size_t q = f; q < ent.size(); ++q) {
                ent[q].unit0 = u;
                u += ent[q].a.n_tiles;
            }
            smp::BatchLaunch<T> d;
            std::memset(&d, 0, sizeof(d));
            d.n_entries = (int32_t)(ent.size() - f);
            d.n_units = u;
            desc.push_back(d);
            lnapp.push_back(napp);
            first.push_back(f);
        }
    }
    const int n_launch = (int)desc.size();
    first.push_back(ent.size());
    // device tables (kept between calls: a captured graph refers to them)
    HIP_TRY(hipSetDevice(b->device));
    const size_t ebytes = sizeof(smp::BatchEntry<T>) * std::max<size_t>(1, ent.size());
    if (ebytes > b->entries_cap || n_launch > b->launch_cap) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        b->g_launches = -1;
        for (void *p : {b->entries, b->launch, (void *)b->counters})
            if (p) (void)hipFree(p);
        b->entries = b->launch = nullptr;
        b->counters = nullptr;
        b->entries_cap = ebytes + ebytes / 4;
        b->launch_cap = n_launch + n_launch / 4 + 1;
        HIP_TRY(hipMalloc(&b->entries, b->entries_cap));
        HIP_TRY(hipMalloc(&b->launch, sizeof(smp::BatchLaunch<T>) * b->launch_cap));
        HIP_TRY(hipMalloc((void **)&b->counters, sizeof(int) * b->launch_cap));
    }
    smp::BatchEntry<T> *dent = (smp::BatchEntry<T> *)b->entries;
    for (int L = 0; L < n_launch; ++L) {
        desc[L].entries = dent + first[L];
        desc[L].counter = b->counters + L;
    }
    HIP_TRY(hipMemcpyAsync(b->entries, ent.data(), sizeof(smp::BatchEntry<T>) * ent.size(),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->launch, desc.data(), sizeof(smp::BatchLaunch<T>) * n_launch,
                           hipMemcpyHostToDevice, b->stream));
    // persistent grid: every CU holds as many workgroups as fit
    // one kernel variant for the whole batch: the problems' cell-wave count
    const int cw = b->hs[0]->cell_waves;
    for (int i = 1; i < n; ++i)
        if (b->hs[i]->cell_waves != cw)
            return fail(SMP_E_INVALID, "batched handles must share cell_waves");
    int ncu = 0, per_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    int per_cu1 = 0;
    const bool pair = b->hs[0]->ts.pair;
    for (int i = 1; i < n; ++i)
        if (b->hs[i]->ts.pair != pair)
            return fail(SMP_E_INVALID, "batched handles must all or none use tiles of two cells per thread");
    HIP_TRY(smp::batch_occupancy<T>(nt, gen, cw, 2, pair, lds, &per_cu));
    HIP_TRY(smp::batch_occupancy<T>(nt, gen, cw, 1, pair, lds, &per_cu1));
    per_cu = std::min(per_cu, per_cu1);
    if (per_cu < 1) return fail(SMP_E_INVALID, "the batched kernel does not fit on a CU");
    const int grid = ncu * per_cu;
    b->grid = grid;
    b->lds = lds;
    const smp::BatchLaunch<T> *dl = (const smp::BatchLaunch<T> *)b->launch;
    auto issue = [&](bool timed) -> int {
        HIP_TRY(smp::launch_fill(b->counters, n_launch, 0, b->stream));
        for (int L = 0; L < n_launch; ++L) {
            if (timed) HIP_TRY(hipEventRecord(b->events[2 + 2 * L], b->stream));
            HIP_TRY(smp::launch_batch<T>(dl + L, nt, gen, cw, lnapp[L], pair, grid, lds, b->stream));
            if (timed) HIP_TRY(hipEventRecord(b->events[3 + 2 * L], b->stream));
        }
        return SMP_OK;
    };
    const bool timed = !use_graph && n_launch <= 4096;
    while ((int)b->events.size() < (timed ? 2 * n_launch + 2 : 2)) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        b->events.push_back(e);
    }
    if (use_graph) {
        const bool same = b->gexec && b->g_launches == n_launch && b->g_grid == grid &&
                          b->g_lds == lds && b->g_nt == nt && b->g_gen == gen && b->g_cw == cw &&
                          b->g_napp == lnapp;
        if (!same) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
            HIP_TRY(hipStreamSynchronize(b->stream));
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = issue(false);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            HIP_TRY(ec);
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            b->g_launches = n_launch;
            b->g_grid = grid;
            b->g_lds = lds;
            b->g_nt = nt;
            b->g_gen = gen;
            b->g_cw = cw;
            b->g_napp = lnapp;
        }
        HIP_TRY(hipEventRecord(b->events[0], b->stream));
        HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
        HIP_TRY(hipEventRecord(b->events[1], b->stream));
    } else {
        HIP_TRY(hipEventRecord(b->events[0], b->stream));
        int rc = issue(timed);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b->events[1], b->stream));
    }
    HIP_TRY(hipEventSynchronize(b->events[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->events[0], b->events[1]));
    b->last_ms = ms;
    double kms = 0;
    if (timed) {
        for (int L = 0; L < n_launch; ++L) {
            float m = 0;
            HIP_TRY(hipEventElapsedTime(&m, b->events[2 + 2 * L], b->events[3 + 2 * L]));
            kms += m;
        }
    } else {
        kms = ms;
    }
    b->last_kernel_ms = kms;
    b->last_launches = n_launch;
    b->last_bytes = bytes;
    for (int i = 0; i < n; ++i) {
        const smp_problem *h = b->hs[i];
        b->last_mcu += (int64_t)h->geo.n_sys * h->geo.n[0] * h->geo.n[1] * h->geo.n[2] * ns[i];
    }
    return SMP_OK;
}

}  // namespace

extern "C" {

int smp_step(smp_problem *h, double dt, int64_t n_steps) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(SMP_E_INVALID, "dt must be positive and finite");
    if (n_steps < 0) return fail(SMP_E_INVALID, "n_steps must be >= 0");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    if (n_steps == 0) return SMP_OK;
    int rc = h->precision == SMP_FP64 ? run_steps<double>(h, dt, n_steps)
                                      : run_steps<float>(h, dt, n_steps);
    if (rc) return rc;
    h->steps_unchecked += n_steps;
    if (h->check_every > 0 && h->steps_unchecked >= h->check_every) {
        bool bad = false;
        rc = scan_state(h, &bad);
        if (rc) return rc;
        if (bad)
            return fail(SMP_E_UNSTABLE, "non-finite values in the state after " +
                                            std::to_string(n_steps) + " steps of dt = " +
                                            std::to_string(dt) +
                                            " (unstable time step: is par.cn too large?)");
    }
    return SMP_OK;
}

int smp_check_finite(smp_problem *h, int32_t *nonfinite) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    bool bad = false;
    int rc = scan_state(h, &bad);
    if (rc) return rc;
    if (nonfinite) *nonfinite = bad ? 1 : 0;
    if (bad) return fail(SMP_E_UNSTABLE, "non-finite values in the state");
    return SMP_OK;
}

int smp_snapshot(smp_problem *h, const int32_t *comps, int32_t n_comps) {
    if (!h || (n_comps > 0 && !comps)) return fail(SMP_E_INVALID, "null argument");
    if (n_comps < 0) return fail(SMP_E_INVALID, "n_comps must be >= 0");
    for (int i = 0; i < n_comps; ++i)
        if (comps[i] < 0 || comps[i] >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    if ((int)h->snap_comps.size() != n_comps) {
        h->dfree(h->snap);
        h->snap = nullptr;
        if (n_comps > 0) {
            int rc = h->dalloc(&h->snap, (size_t)n_comps * h->slab * h->elem);
            if (rc) return rc;
        }
    }
    h->snap_comps.assign(comps, comps + n_comps);
    for (int i = 0; i < n_comps; ++i)
        HIP_TRY(hipMemcpyAsync((char *)h->snap + (size_t)i * h->slab * h->elem,
                               (const char *)h->U + (size_t)comps[i] * h->slab * h->elem,
                               (size_t)h->slab * h->elem, hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

int smp_interp(smp_problem *h, double lambda, double *out, int64_t count) {
    if (!h || !out) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    if (h->snap_comps.empty()) return fail(SMP_E_STATE, "no snapshot taken");
    int64_t total = 0;
    for (int j : h->snap_comps) total += h->numel(j);
    if (count != total) return fail(SMP_E_INVALID, "count does not match the snapshot components");
    HIP_TRY(hipSetDevice(h->device));
    if (h->interp_cap < h->slab) {
        h->dfree(h->interp_buf);
        h->interp_buf = nullptr;
        int rc = h->dalloc(&h->interp_buf, (size_t)h->slab * sizeof(double));
        if (rc) return rc;
        h->interp_cap = h->slab;
    }
    std::vector<double> tmp(h->slab);
    int64_t pos = 0;
    for (size_t i = 0; i < h->snap_comps.size(); ++i) {
        const int j = h->snap_comps[i];
        if (h->precision == SMP_FP64) {
            HIP_TRY(smp::launch_interp<double>((const double *)h->snap + i * h->slab,
                                               (const double *)h->U + (size_t)j * h->slab,
                                               (double *)h->interp_buf, h->slab, lambda, h->stream));
        } else {
            HIP_TRY(smp::launch_interp<float>((const float *)h->snap + i * h->slab,
                                              (const float *)h->U + (size_t)j * h->slab,
                                              (double *)h->interp_buf, h->slab, lambda, h->stream));
        }
        HIP_TRY(hipMemcpyAsync(tmp.data(), h->interp_buf, (size_t)h->slab * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        const int gt[3] = {h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2]};
        h->unpack<double>(gt, tmp.data(), out + pos);
        pos += h->numel(j);
    }
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(out[i]))
            return fail(SMP_E_UNSTABLE, "non-finite values in the interpolated output");
    return SMP_OK;
}

int smp_observables(smp_problem *h, double *sums, double *sq, int32_t *nonfinite) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    const int blocks = 1024;
    double *part = (double *)h->red_buf;
    double *res = part + 3 * blocks;
    double cell = 1;
    for (int e = 0; e < h->geo.ndim; ++e) cell *= h->geo.h[h->geo.axis[e]];
    int bad = 0;
    for (int j = 0; j < h->geo.n_sys; ++j) {
        smp::ReduceGeom g;
        std::memset(&g, 0, sizeof(g));
        for (int e = 0; e < 3; ++e) {
            if (e < h->geo.ndim) {
                const int ax = h->geo.axis[e];
                g.ext[e] = h->geo.ext_of(j, ax);
                g.stag[e] = g.ext[e] != h->geo.n[ax];
            } else {
                g.ext[e] = 1;
                g.stag[e] = 0;
            }
        }
        g.pitch0 = h->pitch0;
        g.ext1 = h->E[1];
        g.cell = cell;
        if (h->precision == SMP_FP64)
            HIP_TRY(smp::launch_reduce<double>((const double *)h->U + (size_t)j * h->slab, g, part,
                                               blocks, h->stream));
        else
            HIP_TRY(smp::launch_reduce<float>((const float *)h->U + (size_t)j * h->slab, g, part,
                                              blocks, h->stream));
        HIP_TRY(smp::launch_reduce_final(part, blocks, res, h->stream));
        double r[3];
        HIP_TRY(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (sums) sums[j] = r[0];
        if (sq) sq[j] = r[1];
        if (r[2] != 0) bad = 1;
    }
    if (nonfinite) *nonfinite = bad;
    return SMP_OK;
}

int smp_moment_norms(smp_problem *h, double *sumsq, double *diffsq) {
    if (!h || !sumsq || !diffsq) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    const int blocks = 1024;
    if (!h->norm_buf) {
        int rc = h->dalloc(&h->norm_buf, sizeof(double) * (4 * (size_t)blocks + 4 * (size_t)h->geo.n_sys));
        if (rc) return rc;
    }
    double *part = (double *)h->norm_buf;
    double *res = part + 4 * blocks;
    std::vector<double> r(4 * (size_t)h->geo.n_sys);
    for (int j = 0; j < h->geo.n_sys; ++j) {
        smp::NormGeom g;
        std::memset(&g, 0, sizeof(g));
        for (int e = 0; e < 3; ++e) {
            if (e < h->geo.ndim) {
                const int ax = h->geo.axis[e];
                g.ext[e] = h->geo.ext_of(j, ax);
                g.rot[e] = h->geo.gt[3 * j + ax] == 2 && h->geo.bc[ax] == 0;
            } else {
                g.ext[e] = 1;
            }
        }
        g.pitch0 = h->pitch0;
        g.ext1 = h->E[1];
        if (h->precision == SMP_FP64)
            HIP_TRY(smp::launch_norm<double>((const double *)h->U + (size_t)j * h->slab, g, part, blocks,
                                             res + 4 * j, h->stream));
        else
            HIP_TRY(smp::launch_norm<float>((const float *)h->U + (size_t)j * h->slab, g, part, blocks,
                                            res + 4 * j, h->stream));
    }
    HIP_TRY(hipMemcpyAsync(r.data(), res, sizeof(double) * r.size(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int j = 0; j < h->geo.n_sys; ++j) {
        sumsq[j] = r[4 * j];
        for (int e = 0; e < 3; ++e) diffsq[3 * j + e] = r[4 * j + 1 + e];
    }
    return SMP_OK;
}

int smp_probe(smp_problem *h, const double *lo, const double *point, double *out) {
    if (!h || !lo || !point || !out) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    std::vector<smp::ProbeComp> pc;
    probe_table(h, lo, point, pc);
    void *dpc = nullptr, *dout = nullptr;
    int rc = h->dalloc(&dpc, sizeof(smp::ProbeComp) * pc.size());
    if (!rc) rc = h->dalloc(&dout, sizeof(double) * pc.size());
    if (!rc) rc = probe_run(h, pc, (smp::ProbeComp *)dpc, (double *)dout, out);
    h->dfree(dpc);
    h->dfree(dout);
    return rc;
}

int smp_probe_noalloc(smp_problem *h, const double *lo, const double *point, double *out) {
    if (!h || !lo || !point || !out) return fail(SMP_E_INVALID, "null argument");
    if (h->geo.n_sys > SMP_PROBE_NOALLOC_MAX_COMPS)
        return fail(SMP_E_INVALID, "smp_probe_noalloc takes at most " +
                                       std::to_string(SMP_PROBE_NOALLOC_MAX_COMPS) +
                                       " components (this problem has " +
                                       std::to_string(h->geo.n_sys) + ")");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    char *base = nullptr;
    int rc = probe_scratch(h->device, &base);
    if (rc) return rc;
    std::vector<smp::ProbeComp> pc;
    probe_table(h, lo, point, pc);
    return probe_run(h, pc, (smp::ProbeComp *)base,
                     (double *)(base + SMP_PROBE_NOALLOC_MAX_COMPS * sizeof(smp::ProbeComp)), out);
}

int smp_last_step_timing(const smp_problem *h, double *total_ms, double *substep_ms,
                         int64_t *substep_launches, double *substep_bytes) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (substep_bytes) *substep_bytes = h->last_bytes;
    if (total_ms) *total_ms = h->last_total_ms;
    if (substep_ms) *substep_ms = h->last_substep_ms;
    if (substep_launches) *substep_launches = h->last_launches;
    return SMP_OK;
}

void *smp_stream(const smp_problem *h) { return h ? (void *)h->stream : nullptr; }

int64_t smp_device_bytes(const smp_problem *h) { return h ? h->dev_bytes : 0; }

int smp_plan_stats(const smp_problem *h, int32_t kind, int32_t *n_groups, int32_t *n_slots,
                   int32_t *n_loads, int32_t *n_sources) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (kind != 0 && kind != 1) return fail(SMP_E_INVALID, "kind must be 0 or 1");
    const smp::Plan &p = h->plan[kind];
    if (n_groups) *n_groups = (int32_t)p.groups.size();
    if (n_slots) *n_slots = p.n_slots;
    if (n_loads) *n_loads = (int32_t)p.loads.size();
    if (n_sources) *n_sources = p.n_sources;
    return SMP_OK;
}

int smp_set_schedule(smp_problem *h, const int32_t *tile, int32_t blocks_per_cu,
                     int32_t cell_waves) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (blocks_per_cu < 0 || blocks_per_cu > 8) return fail(SMP_E_INVALID, "blocks_per_cu must be 0..8");
    if (cell_waves != 0 && cell_waves != 4 && cell_waves != 8)
        return fail(SMP_E_INVALID, "cell_waves must be 0, 4 or 8");
    if (!h->host_only) {
        HIP_TRY(hipSetDevice(h->device));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    smp::TileShape ts;
    if (tile && (tile[0] || tile[1] || tile[2])) {
        std::string err = smp::make_tile(h->geo, std::max(1, tile[0]), std::max(1, tile[1]),
                                         std::max(1, tile[2]), ts);
        if (!err.empty()) return fail(SMP_E_INVALID, err);
    } else {
        ts = smp::choose_tile(h->geo, h->elem);
    }
    if (ts.pair && h->precision != SMP_FP32)
        return fail(SMP_E_INVALID, "tiles of more than 256 cells (two cells per thread) need the "
                                   "FP32 state");
    if ((ts.slot_elems * (h->elem / 4) + 63) / 64 > smp::CPS_MAX)
        return fail(SMP_E_INVALID, "a tile with its halo exceeds 4 KB per component");
    const smp::TileShape old_ts = h->ts;
    const int old_blocks = h->blocks_per_cu;
    h->ts = ts;
    if (blocks_per_cu) h->blocks_per_cu = blocks_per_cu;
    for (int kind = 0; kind < 2; ++kind) {
        int rc = make_plan(h, kind, h->ncf_cap[kind]);
        if (rc) {
            // keep the handle usable with its previous schedule
            const std::string msg = g_err;
            h->ts = old_ts;
            h->blocks_per_cu = old_blocks;
            for (int k2 = 0; k2 < 2; ++k2) make_plan(h, k2, h->ncf_cap[k2]);
            h->coef_dirty = true;
            return fail(rc, msg);
        }
    }
    h->n_tiles = 1;
    for (int e = 0; e < 3; ++e) {
        h->ntile[e] = (h->E[e] + h->ts.tile[e] - 1) / h->ts.tile[e];
        h->n_tiles *= h->ntile[e];
    }
    if (cell_waves) h->cell_waves = cell_waves;
    h->coef_dirty = true;
    return SMP_OK;
}

int smp_get_schedule(const smp_problem *h, int32_t *tile, int32_t *blocks_per_cu,
                     int32_t *cell_waves) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (tile)
        for (int e = 0; e < 3; ++e) tile[e] = h->ts.tile[e];
    if (blocks_per_cu) *blocks_per_cu = h->blocks_per_cu;
    if (cell_waves) *cell_waves = h->cell_waves;
    return SMP_OK;
}

int smp_checkpoint_save(smp_problem *h, int32_t slot) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    void *dst = nullptr;
    int rc = checkpoint_buffer(h, slot, &dst);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(dst, h->U, (size_t)h->slab * h->elem * h->geo.n_sys,
                           hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

int smp_checkpoint_load(smp_problem *h, int32_t slot) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    REQUIRE_DEVICE(h);
    if (slot < 0 || slot >= (int)h->ckpt.size() || !h->ckpt[slot])
        return fail(SMP_E_STATE, "checkpoint slot " + std::to_string(slot) + " was never saved");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemcpyAsync(h->U, h->ckpt[slot], (size_t)h->slab * h->elem * h->geo.n_sys,
                           hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

int smp_checkpoint_free(smp_problem *h, int32_t slot) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (!h->host_only) {
        HIP_TRY(hipSetDevice(h->device));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    for (int s = 0; s < (int)h->ckpt.size(); ++s) {
        if (slot >= 0 && s != slot) continue;
        h->dfree(h->ckpt[s]);
        h->ckpt[s] = nullptr;
    }
    return SMP_OK;
}

int smp_batch_create(smp_problem *const *handles, int32_t n, smp_batch **out) {
    g_err.clear();
    if (!handles || !out || n < 1) return fail(SMP_E_INVALID, "a batch needs at least one handle");
    *out = nullptr;
    std::unique_ptr<smp_batch> b(new smp_batch());
    for (int i = 0; i < n; ++i) {
        smp_problem *h = handles[i];
        if (!h) return fail(SMP_E_INVALID, "null handle in the batch");
        if (h->host_only) return fail(SMP_E_STATE, "handle was created without a device (device = -1)");
        if (i > 0 && (h->device != handles[0]->device || h->precision != handles[0]->precision))
            return fail(SMP_E_INVALID, "batched handles must share device and precision");
        if (h->f_position == SMP_FILTER_FULL)
            return fail(SMP_E_INVALID, "batched stepping supports no filter or the substep filter");
        for (int k = 0; k < i; ++k)
            if (handles[k] == h) return fail(SMP_E_INVALID, "a handle appears twice in the batch");
        b->hs.push_back(h);
    }
    b->device = handles[0]->device;
    b->precision = handles[0]->precision;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    *out = b.release();
    return SMP_OK;
}

int smp_batch_destroy(smp_batch *b) {
    delete b;
    return SMP_OK;
}

int smp_batch_run(smp_batch *b, const double *dt, const int64_t *n_steps, int32_t use_graph) {
    if (!b || !dt || !n_steps) return fail(SMP_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(b->device));
    return b->precision == SMP_FP64 ? batch_run<double>(b, dt, n_steps, use_graph)
                                    : batch_run<float>(b, dt, n_steps, use_graph);
}

int smp_batch_timing(const smp_batch *b, double *total_ms, double *kernel_ms, int64_t *launches,
                     double *bytes, int64_t *mcu, int32_t *grid) {
    if (!b) return fail(SMP_E_INVALID, "null batch");
    if (total_ms) *total_ms = b->last_ms;
    if (kernel_ms) *kernel_ms = b->last_kernel_ms;
    if (launches) *launches = b->last_launches;
    if (bytes) *bytes = b->last_bytes;
    if (mcu) *mcu = b->last_mcu;
    if (grid) *grid = b->grid;
    return SMP_OK;
}

int smp_adjoint_reset(smp_problem *h) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    int rc = adjoint_buffers(h);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(h->lam, 0, (size_t)h->slab * h->elem * h->geo.n_sys, h->stream));
    HIP_TRY(hipMemsetAsync(h->G, 0, (size_t)h->slab * 8 * 16, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SMP_OK;
}

int smp_adjoint_seed(smp_problem *h, int32_t j, const double *src, int64_t count) {
    if (!h || !src) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    if (j < 0 || j >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    if (count != h->numel(j)) return fail(SMP_E_INVALID, "count does not match the component shape");
    HIP_TRY(hipSetDevice(h->device));
    int rc = adjoint_buffers(h);
    if (rc) return rc;
    const int gt[3] = {h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2]};
    char *dst = (char *)h->lam + (size_t)j * h->slab * h->elem;
    if (h->precision == SMP_FP64) {
        std::vector<double> tmp(h->slab);
        h->pack<double>(gt, src, tmp.data());
        HIP_TRY(hipMemcpy(dst, tmp.data(), (size_t)h->slab * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(h->slab);
        h->pack<float>(gt, src, tmp.data());
        HIP_TRY(hipMemcpy(dst, tmp.data(), (size_t)h->slab * 4, hipMemcpyHostToDevice));
    }
    return SMP_OK;
}

int smp_adjoint(smp_problem *h, double dt, int64_t n_steps, int32_t checkpoint_every) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(SMP_E_INVALID, "dt must be positive and finite");
    if (n_steps < 1) return fail(SMP_E_INVALID, "n_steps must be >= 1");
    if (checkpoint_every < 1) return fail(SMP_E_INVALID, "checkpoint_every must be >= 1");
    REQUIRE_DEVICE(h);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int every = (int)std::min<int64_t>(checkpoint_every, n_steps);
    return h->precision == SMP_FP64 ? run_adjoint<double>(h, dt, n_steps, every)
                                    : run_adjoint<float>(h, dt, n_steps, every);
}

int smp_adjoint_get(smp_problem *h, int32_t what, int32_t j, double *dst, int64_t count) {
    if (!h || !dst) return fail(SMP_E_INVALID, "null argument");
    REQUIRE_DEVICE(h);
    if (what < 0 || what > 2)
        return fail(SMP_E_INVALID, "what must be 0 (adjoint state), 1 (dJ/dsigma_a) or 2 (dJ/dsigma_s0)");
    if (j < 0 || j >= h->geo.n_sys) return fail(SMP_E_INVALID, "component out of range");
    if (count != h->numel(j)) return fail(SMP_E_INVALID, "count does not match the component shape");
    if (!h->lam) return fail(SMP_E_STATE, "no adjoint was run or seeded");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int gt[3] = {h->geo.gt[3 * j], h->geo.gt[3 * j + 1], h->geo.gt[3 * j + 2]};
    if (what >= 1) {
        const int code = (gt[0] - 1) * 4 + (gt[1] - 1) * 2 + (gt[2] - 1) + (what == 2 ? 8 : 0);
        std::vector<double> tmp(h->slab);
        HIP_TRY(hipMemcpy(tmp.data(), (const char *)h->G + (size_t)code * h->slab * 8,
                          (size_t)h->slab * 8, hipMemcpyDeviceToHost));
        h->unpack<double>(gt, tmp.data(), dst);
    } else if (h->precision == SMP_FP64) {
        std::vector<double> tmp(h->slab);
        HIP_TRY(hipMemcpy(tmp.data(), (const char *)h->lam + (size_t)j * h->slab * 8,
                          (size_t)h->slab * 8, hipMemcpyDeviceToHost));
        h->unpack<double>(gt, tmp.data(), dst);
    } else {
        std::vector<float> tmp(h->slab);
        HIP_TRY(hipMemcpy(tmp.data(), (const char *)h->lam + (size_t)j * h->slab * 4,
                          (size_t)h->slab * 4, hipMemcpyDeviceToHost));
        h->unpack<float>(gt, tmp.data(), dst);
    }
    return SMP_OK;
}

int smp_adjoint_timing(const smp_problem *h, double *forward_ms, double *reverse_ms) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (forward_ms) *forward_ms = h->adj_fwd_ms;
    if (reverse_ms) *reverse_ms = h->adj_rev_ms;
    return SMP_OK;
}

int smp_adjoint_set_memory(smp_problem *h, int64_t bytes) {
    if (!h) return fail(SMP_E_INVALID, "null handle");
    if (bytes < 0) return fail(SMP_E_INVALID, "bytes must be >= 0");
    h->adj_budget = bytes;
    return SMP_OK;
}

static void adjoint_stats_of(const smp::AdjSchedule &sc, int64_t bytes, int64_t stats[8]) {
    stats[0] = sc.segments;
    stats[1] = sc.seg_steps;
    stats[2] = sc.tail;
    stats[3] = sc.slots;
    stats[4] = sc.recomputed;
    stats[5] = sc.extra_advances;
    stats[6] = sc.buffers;
    stats[7] = bytes;
}

int smp_adjoint_stats(const smp_problem *h, int64_t stats[8]) {
    if (!h || !stats) return fail(SMP_E_INVALID, "null argument");
    adjoint_stats_of(h->adj_sched, h->adj_peak_bytes, stats);
    return SMP_OK;
}

int smp_adjoint_schedule(int64_t n_steps, int32_t every, int64_t buffers, int32_t rec_factor,
                         int64_t stats[8], int32_t *actions, int64_t cap, int64_t *n_fwd,
                         int64_t *n_actions) {
    if (!stats || !n_fwd || !n_actions) return fail(SMP_E_INVALID, "null argument");
    smp::AdjSchedule sc;
    const std::string err = smp::plan_adjoint(n_steps, every, buffers, rec_factor, sc);
    if (!err.empty()) return fail(buffers >= 0 ? SMP_E_OOM : SMP_E_INVALID, err);
    adjoint_stats_of(sc, 0, stats);
    *n_fwd = (int64_t)sc.fwd.size();
    *n_actions = (int64_t)(sc.fwd.size() + sc.rev.size());
    if (actions) {
        if (cap < *n_actions) return fail(SMP_E_INVALID, "actions buffer too small");
        int64_t q = 0;
        for (const auto *v : {&sc.fwd, &sc.rev})
            for (const smp::AdjAction &a : *v) {
                const int32_t row[6] = {a.kind, a.s0, a.s1, a.src, a.dst, a.range};
                std::memcpy(actions + 6 * q, row, sizeof(row));
                ++q;
            }
    }
    return SMP_OK;
}

}  // extern "C"


