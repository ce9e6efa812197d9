// Float64 reductions over a resident scene: the shutoff energies and the plane-monitor flux.
//
// Both are two-stage and deterministic: every block writes a partial sum (wave shuffles, then
// LDS across the block's waves) and one block adds the partials in block order.  No atomics, so
// a scene gives the same bits on every call.

#include "oem_internal.h"

#include <mutex>

#define OEM_RED_BLOCKS 2048
#define OEM_RED_THREADS 256

namespace {

struct EnergyDev {
    const double *dl[3], *dd[3];   // primal and dual cell lengths per axis
    const float *eps;              // per class (n_classes * 4) or dense (3 * ncell)
    int eps_per_class;
    const uint16_t *cls;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum NV doubles over the block; thread 0 receives the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double s[OEM_RED_THREADS / 64][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) s[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0.0;
            for (int u = 0; u < OEM_RED_THREADS / 64; ++u) t += s[u][q];
            v[q] = t;
        }
}

// Terms rounded as separate float32 operations (no fused multiply-add), so they are the exact
// float32 values w * a * a and w * (re * re + im * im).
template <typename F> __device__ __forceinline__ float wsq(float w, F x);
template <> __device__ __forceinline__ float wsq<float>(float w, float x) {
    return __fmul_rn(__fmul_rn(w, x), x);
}
template <> __device__ __forceinline__ float wsq<cf2>(float w, cf2 x) {
    return __fmul_rn(w, __fadd_rn(__fmul_rn(x.r, x.r), __fmul_rn(x.i, x.i)));
}

// partial[3 * block + q]: q = 0 sum V|E|^2, 1 sum V eps |E|^2, 2 sum V |H|^2, with V the E
// component's Yee dual volume rounded to float32 and each term formed in float32.
template <typename F>
__global__ void __launch_bounds__(OEM_RED_THREADS)
k_energy(const F *__restrict__ E, const F *__restrict__ H, EnergyDev d, int nx, int ny, int nz,
         double *__restrict__ partial)
{
    const size_t ncell = (size_t)nx * ny * nz;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < ncell;
         id += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(id % nz);
        const size_t r = id / nz;
        const int j = (int)(r % ny), i = (int)(r / ny);
        const double xp = d.dl[0][i], yp = d.dl[1][j], zp = d.dl[2][k];
        const double xd = d.dd[0][i], yd = d.dd[1][j], zd = d.dd[2][k];
        const float v[3] = {(float)(xp * yd * zd), (float)(xd * yp * zd), (float)(xd * yd * zp)};
        float ep[3];
        if (d.eps_per_class) {
            const float *row = d.eps + 4 * (size_t)d.cls[id];
            ep[0] = row[0]; ep[1] = row[1]; ep[2] = row[2];
        } else {
            ep[0] = d.eps[id]; ep[1] = d.eps[ncell + id]; ep[2] = d.eps[2 * ncell + id];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const F e = E[(size_t)c * ncell + id], h = H[(size_t)c * ncell + id];
            acc[0] += (double)wsq<F>(v[c], e);
            acc[1] += (double)wsq<F>(__fmul_rn(v[c], ep[c]), e);
            acc[2] += (double)wsq<F>(v[c], h);
        }
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partial[3 * (size_t)blockIdx.x + q] = acc[q];
}

// out[q] = sum over blocks of partial[nv * b + q], in block order (one thread per q).
__global__ void k_sum_partials(const double *__restrict__ partial, int nblk, int nv,
                               double *__restrict__ out)
{
    const int q = threadIdx.x;
    if (q >= nv) return;
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += partial[(size_t)nv * b + q];
    out[q] = t;
}

// Plane flux terms per frequency f (grid.y): t1 = sum w1 Re(E_t1 conj H_t2),
// t2 = sum w2 Re(E_t2 conj H_t1) over the plane; partial[(f * nblk + b) * 2 + {0, 1}].
__global__ void __launch_bounds__(OEM_RED_THREADS)
k_flux_partial(const double *__restrict__ re, const double *__restrict__ im,
               const double *__restrict__ w1, const double *__restrict__ w2, int nf, int npts,
               double *__restrict__ partial)
{
    const int f = blockIdx.y;
    const size_t plane = (size_t)nf * npts;
    const double *e1r = re + (size_t)f * npts, *e1i = im + (size_t)f * npts;
    const double *e2r = e1r + plane, *e2i = e1i + plane;
    const double *h1r = e1r + 2 * plane, *h1i = e1i + 2 * plane;
    const double *h2r = e1r + 3 * plane, *h2i = e1i + 3 * plane;
    double acc[2] = {0.0, 0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += gridDim.x * blockDim.x) {
        acc[0] += (e1r[q] * h2r[q] + e1i[q] * h2i[q]) * w1[q];
        acc[1] += (e2r[q] * h1r[q] + e2i[q] * h1i[q]) * w2[q];
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partial[((size_t)f * gridDim.x + blockIdx.x) * 2] = acc[0];
        partial[((size_t)f * gridDim.x + blockIdx.x) * 2 + 1] = acc[1];
    }
}

// out[2 f + t] = sum over blocks b of partial[(f * nblk + b) * 2 + t]
__global__ void k_sum_flux(const double *__restrict__ partial, int nblk, int nf,
                           double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nf) return;
    const int f = t >> 1, c = t & 1;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[((size_t)f * nblk + b) * 2 + c];
    out[t] = s;
}

// Weighted complex sums of a plane monitor's phasors: for set s and frequency f (grid.y =
// s * nf + f), sum over points q and components c of D_c[f][q] * G[s][c][f][q] (complex), with
// D the monitor's float64 accumulators and G (re, im) pairs; partial[(y * nblk + b) * 2 + {re, im}].
__global__ void __launch_bounds__(OEM_RED_THREADS)
k_overlap_partial(const double *__restrict__ re, const double *__restrict__ im,
                  const double *__restrict__ g, int nf, int npts, double *__restrict__ partial)
{
    const int y = blockIdx.y, f = y % nf, s = y / nf;
    const size_t plane = (size_t)nf * npts;
    double acc[2] = {0.0, 0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t o = c * plane + (size_t)f * npts + q;
            const double dr = re[o], di = im[o];
            const double *w = g + 2 * (((size_t)s * 4 + c) * plane + (size_t)f * npts + q);
            acc[0] += dr * w[0] - di * w[1];
            acc[1] += dr * w[1] + di * w[0];
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partial[((size_t)y * gridDim.x + blockIdx.x) * 2] = acc[0];
        partial[((size_t)y * gridDim.x + blockIdx.x) * 2 + 1] = acc[1];
    }
}

}  // namespace

// Reduction state kept on the context: tables for the energies, flux weights per monitor and
// a scratch area for the partial sums.
struct oem_reduce_state {
    EnergyDev ed{};
    bool energy_ready = false;
    std::vector<std::pair<double *, double *>> flux_w;   // per monitor (w1, w2) or nulls
    std::vector<std::pair<double *, int>> over_w;        // per monitor (G, n_sets) or null
    double *partial = nullptr;                           // 3 * OEM_RED_BLOCKS doubles at least
    size_t partial_n = 0;
    double *out = nullptr;                               // small result buffer
};

namespace oemi {
void reduce_state_free(oem_reduce_state *r) { delete r; }
}

static int red_state(oem_ctx *ctx, oem_reduce_state **out) {
    if (!ctx->red) ctx->red = new oem_reduce_state();
    *out = ctx->red;
    return OEM_OK;
}

static int red_scratch(oem_ctx *ctx, oem_reduce_state *r, size_t need) {
    if (r->partial_n < need) {
        TRY(ctx->alloc(&r->partial, need, false));   // released with the context
        r->partial_n = need;
    }
    if (!r->out) TRY(ctx->alloc(&r->out, 2 * 65536, false));
    return OEM_OK;
}

extern "C" int oem_set_energy_tables(oem_ctx *ctx, const double *dl_x, const double *dl_y,
                                     const double *dl_z, const double *dd_x, const double *dd_y,
                                     const double *dd_z, const float *eps,
                                     int32_t eps_per_class) {
    CHECK_CTX(ctx);
    if (!ctx->mat_set) return fail(OEM_ESTATE, "materials not set");
    const double *dl[3] = {dl_x, dl_y, dl_z}, *dd[3] = {dd_x, dd_y, dd_z};
    for (int a = 0; a < 3; ++a)
        if (!dl[a] || !dd[a]) return fail(OEM_EINVAL, "axis %d: null lengths", a);
    if (!eps) return fail(OEM_EINVAL, "null eps");
    if (eps_per_class && ctx->mat.mode != 0)
        return fail(OEM_EINVAL, "per-class eps needs the class LUT materials");
    oem_reduce_state *r;
    TRY(red_state(ctx, &r));
    const int n[3] = {ctx->dims.nx, ctx->dims.ny, ctx->dims.nz};
    for (int a = 0; a < 3; ++a) {
        double *p, *q;
        TRY(ctx->upload(&p, dl[a], (size_t)n[a]));
        TRY(ctx->upload(&q, dd[a], (size_t)n[a]));
        r->ed.dl[a] = p;
        r->ed.dd[a] = q;
    }
    float *pe;
    const size_t ne = eps_per_class ? 4 * (size_t)ctx->mat.n_classes : 3 * ctx->ncell;
    TRY(ctx->upload(&pe, eps, ne));
    r->ed.eps = pe;
    r->ed.eps_per_class = eps_per_class ? 1 : 0;
    r->ed.cls = ctx->mat.cls;
    r->energy_ready = true;
    return OEM_OK;
}

extern "C" int oem_energy(oem_ctx *ctx, double *out) {
    CHECK_CTX(ctx);
    if (!out) return fail(OEM_EINVAL, "null output");
    if (!ctx->red || !ctx->red->energy_ready) return fail(OEM_ESTATE, "energy tables not set");
    if (ctx->attached) return fail(OEM_ESTATE, "context is stepped by its pool");
    oem_reduce_state *r = ctx->red;
    TRY(red_scratch(ctx, r, 3 * (size_t)OEM_RED_BLOCKS));
    const dim3 g(OEM_RED_BLOCKS), b(OEM_RED_THREADS);
    if (ctx->cplx)
        hipLaunchKernelGGL((k_energy<cf2>), g, b, 0, ctx->stream, (const cf2 *)ctx->cur_E(),
                           (const cf2 *)ctx->cur_H(), r->ed, ctx->dims.nx, ctx->dims.ny, ctx->dims.nz,
                           r->partial);
    else
        hipLaunchKernelGGL((k_energy<float>), g, b, 0, ctx->stream, (const float *)ctx->cur_E(),
                           (const float *)ctx->cur_H(), r->ed, ctx->dims.nx, ctx->dims.ny, ctx->dims.nz,
                           r->partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, ctx->stream, r->partial,
                       OEM_RED_BLOCKS, 3, r->out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, r->out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_set_flux_weights(oem_ctx *ctx, int32_t mon_id, const double *w1,
                                    const double *w2) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const MonDev &m = ctx->mons[mon_id].d;
    if (m.kind != 0) return fail(OEM_EINVAL, "monitor %d is not a plane", mon_id);
    if (!w1 || !w2) return fail(OEM_EINVAL, "null weights");
    oem_reduce_state *r;
    TRY(red_state(ctx, &r));
    if (r->flux_w.size() < ctx->mons.size()) r->flux_w.resize(ctx->mons.size(), {nullptr, nullptr});
    double *p, *q;
    TRY(ctx->upload(&p, w1, (size_t)m.npts));
    TRY(ctx->upload(&q, w2, (size_t)m.npts));
    r->flux_w[mon_id] = {p, q};
    return OEM_OK;
}

extern "C" int oem_monitor_flux(oem_ctx *ctx, int32_t mon_id, double *out, size_t count) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const MonDev &m = ctx->mons[mon_id].d;
    if (!out || count != 2 * (size_t)m.nf)
        return fail(OEM_EINVAL, "monitor %d: output holds 2*nf = %d doubles", mon_id, 2 * m.nf);
    oem_reduce_state *r = ctx->red;
    if (!r || (int)r->flux_w.size() <= mon_id || !r->flux_w[mon_id].first)
        return fail(OEM_ESTATE, "flux weights of monitor %d not set", mon_id);
    if (ctx->pool) TRY(oemi::pool_wait_done(ctx));
    const int nblk = std::max(1, std::min(64, (m.npts + OEM_RED_THREADS - 1) / OEM_RED_THREADS));
    TRY(red_scratch(ctx, r, std::max((size_t)3 * OEM_RED_BLOCKS, (size_t)2 * nblk * m.nf)));
    hipLaunchKernelGGL(k_flux_partial, dim3(nblk, m.nf), dim3(OEM_RED_THREADS), 0, ctx->stream,
                       m.re, m.im, r->flux_w[mon_id].first, r->flux_w[mon_id].second, m.nf,
                       m.npts, r->partial);
    hipLaunchKernelGGL(k_sum_flux, dim3((2 * m.nf + 255) / 256), dim3(256), 0, ctx->stream,
                       r->partial, nblk, m.nf, r->out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, r->out, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_set_overlap_weights(oem_ctx *ctx, int32_t mon_id, int32_t n_sets,
                                       const double *g) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const MonDev &m = ctx->mons[mon_id].d;
    if (m.kind != 0) return fail(OEM_EINVAL, "monitor %d is not a plane", mon_id);
    if (!g || n_sets < 1 || n_sets > 16) return fail(OEM_EINVAL, "bad overlap weights (n_sets %d)", n_sets);
    oem_reduce_state *r;
    TRY(red_state(ctx, &r));
    if (r->over_w.size() < ctx->mons.size()) r->over_w.resize(ctx->mons.size(), {nullptr, 0});
    double *p;
    TRY(ctx->upload(&p, g, (size_t)n_sets * 4 * m.nf * m.npts * 2));
    r->over_w[mon_id] = {p, n_sets};
    return OEM_OK;
}

extern "C" int oem_monitor_overlap(oem_ctx *ctx, int32_t mon_id, double *out, size_t count) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const MonDev &m = ctx->mons[mon_id].d;
    oem_reduce_state *r = ctx->red;
    if (!r || (int)r->over_w.size() <= mon_id || !r->over_w[mon_id].first)
        return fail(OEM_ESTATE, "overlap weights of monitor %d not set", mon_id);
    const int ns = r->over_w[mon_id].second;
    if (!out || count != 2 * (size_t)ns * m.nf)
        return fail(OEM_EINVAL, "monitor %d: output holds 2*n_sets*nf = %d doubles", mon_id, 2 * ns * m.nf);
    if (ctx->pool) TRY(oemi::pool_wait_done(ctx));
    const int nblk = std::max(1, std::min(64, (m.npts + OEM_RED_THREADS - 1) / OEM_RED_THREADS));
    const int ny = ns * m.nf;
    TRY(red_scratch(ctx, r, std::max((size_t)3 * OEM_RED_BLOCKS, (size_t)2 * nblk * ny)));
    hipLaunchKernelGGL(k_overlap_partial, dim3(nblk, ny), dim3(OEM_RED_THREADS), 0, ctx->stream,
                       m.re, m.im, r->over_w[mon_id].first, m.nf, m.npts, r->partial);
    hipLaunchKernelGGL(k_sum_flux, dim3((2 * ny + 255) / 256), dim3(256), 0, ctx->stream,
                       r->partial, nblk, ny, r->out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, r->out, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

// ------------------------------------------------------------------------------------------
// Plane flux without allocation.  The partial sums (64 blocks x nf x 2 doubles at most) and the
// 2 nf results of oem_monitor_flux_noalloc live in `scratch`, one file-scope device array per
// GPU, instead of the context's lazily allocated reduction buffers, so the call never reaches
// hipMalloc.  Its device address is looked up once per GPU, with that GPU current, on the first
// call there; host code only ever passes that address, never the array's name.  Calls on one
// GPU share the array: a host mutex per GPU holds it for the whole call (launches, copy, stream
// synchronisation), so calls from several threads or contexts on one GPU run one after another;
// calls issued outside this entry point that overlap it on the same GPU would race.
#define OEM_NOALLOC_MAX_NF 1024
#define OEM_NOALLOC_BLOCKS 64
#define OEM_MAX_DEVICES 64

static __device__ __align__(256) double
    scratch[2 * OEM_NOALLOC_BLOCKS * OEM_NOALLOC_MAX_NF + 2 * OEM_NOALLOC_MAX_NF];

static std::mutex g_scratch_mu[OEM_MAX_DEVICES];
static double *g_scratch_ptr[OEM_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (the current device), looked up on first use.
static int scratch_ptr(int device, double **out) {
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OEM_EHIP, "device %d: address of the flux scratch array: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return OEM_OK;
}

extern "C" int oem_monitor_flux_noalloc(oem_ctx *ctx, int32_t mon_id, double *out, size_t count) {
    CHECK_CTX(ctx);
    if (ctx->device < 0 || ctx->device >= OEM_MAX_DEVICES)
        return fail(OEM_EINVAL, "device %d beyond the %d the scratch table covers", ctx->device,
                    OEM_MAX_DEVICES);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const MonDev &m = ctx->mons[mon_id].d;
    if (!out || count != 2 * (size_t)m.nf)
        return fail(OEM_EINVAL, "monitor %d: output holds 2*nf = %d doubles", mon_id, 2 * m.nf);
    if (m.nf > OEM_NOALLOC_MAX_NF)
        return fail(OEM_EINVAL, "monitor %d: %d frequencies, the allocation-free flux takes at most %d",
                    mon_id, m.nf, OEM_NOALLOC_MAX_NF);
    oem_reduce_state *r = ctx->red;
    if (!r || (int)r->flux_w.size() <= mon_id || !r->flux_w[mon_id].first)
        return fail(OEM_ESTATE, "flux weights of monitor %d not set", mon_id);
    if (ctx->pool) TRY(oemi::pool_wait_done(ctx));
    std::lock_guard<std::mutex> hold(g_scratch_mu[ctx->device]);
    double *base;
    TRY(scratch_ptr(ctx->device, &base));
    double *partial = base;
    double *res = base + 2 * OEM_NOALLOC_BLOCKS * OEM_NOALLOC_MAX_NF;
    // the same launch shape as oem_monitor_flux, so both give the same bits
    const int nblk = std::max(1, std::min(OEM_NOALLOC_BLOCKS, (m.npts + OEM_RED_THREADS - 1) / OEM_RED_THREADS));
    hipLaunchKernelGGL(k_flux_partial, dim3(nblk, m.nf), dim3(OEM_RED_THREADS), 0, ctx->stream,
                       m.re, m.im, r->flux_w[mon_id].first, r->flux_w[mon_id].second, m.nf,
                       m.npts, partial);
    hipLaunchKernelGGL(k_sum_flux, dim3((2 * m.nf + 255) / 256), dim3(256), 0, ctx->stream,
                       partial, nblk, m.nf, res);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, res, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}
