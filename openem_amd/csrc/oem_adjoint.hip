// Discrete adjoint of the time step (real fields, no dispersion): checkpoints, recomputation of
// a segment with E recorded at the design cells, and the backward sweep.
//
// The forward step is linear in (E, H, psi_H, psi_E) plus source terms; with
//   H' = H - ch (s1 - s2),  s = (1/kappa + a) D_fwd E + b psi_H,  psi_H' = b psi_H + a D_fwd E
//   E' = pec (ca E + cb (s1 - s2)),  s = (1/kappa + a) D_bwd H' + b psi_E,  psi_E' = ...
// the backward sweep applies, per step n from the last: the objective's gradient with respect to
// E^{n+1} and H^{n+1/2} (monitor samples), the eps gradient lambda_E . (ca' E^n + cb' S^n), the
// transpose of the E update (gather from the forward neighbour along each axis, then the
// cell-local scaling), and the transpose of the H update (gather from the backward neighbour).
// Each transpose is two kernels so that no kernel reads a value another thread overwrites.
//
// Each term of an update: (target component c, source component b, derivative axis a, sign,
// psi slot).  psi slot = off[2 a + m] of the slab layout (x: Hy/Ey 0, Hz/Ez 1; y: Hx/Ex 0,
// Hz/Ez 1; z: Hx/Ex 0, Hy/Ey 1).

#include "oem_internal.h"

#define OEM_ADJ_THREADS 256
#define OEM_ADJ_BLOCKS 4096

namespace {

struct Term { int c, b, a, sgn, slot; };
// H_c and E_c share the term structure (curl): c = 0: +d_y(F_z) - d_z(F_y), ...
__constant__ Term kTerms[6] = {
    {0, 2, 1, +1, 2}, {0, 1, 2, -1, 4},     // x: +y(z) [y-slab, Hx/Ex], -z(y) [z-slab, Hx/Ex]
    {1, 0, 2, +1, 5}, {1, 2, 0, -1, 0},     // y: +z(x) [z-slab, Hy/Ey], -x(z) [x-slab, Hy/Ey]
    {2, 1, 0, +1, 1}, {2, 0, 1, -1, 3},     // z: +x(y) [x-slab, Hz/Ez], -y(x) [y-slab, Hz/Ez]
};

struct AdjDev {
    float *lE, *lH;              // 3 * ncell each
    float *lpe, *lph;            // psi adjoints, slab layout of PsiDev
    PsiDev psl;                  // offsets (pointers unused)
    GridDev g;
    MatDev mat;
    float ch;
    size_t ncell;
};

__device__ __forceinline__ void cell_ijk(size_t id, int ny, int nz, int &i, int &j, int &k) {
    k = (int)(id % nz);
    const size_t r = id / nz;
    j = (int)(r % ny);
    i = (int)(r / ny);
}

__device__ __forceinline__ size_t shift(size_t id, int a, int from, int to, int ny, int nz) {
    const size_t st = a == 0 ? (size_t)ny * nz : (a == 1 ? (size_t)nz : 1);
    return id + (size_t)((long long)to - from) * st;
}

// psi index of cell (i, j, k) in the slab of axis a (or -1 outside the slab)
__device__ __forceinline__ long long psi_at(const AdjDev &d, int slot, int i, int j, int k) {
    const GridDev &g = d.g;
    const int a = slot >> 1;
    const AxisDev &A = g.ax[a];
    const int n = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
    const int p = a == 0 ? i : (a == 1 ? j : k);
    const int s = slab_of(p, n, A.lo, A.hi);
    if (s < 0) return -1;
    const int w = A.lo + A.hi;
    size_t o;
    if (a == 0) o = ((size_t)s * g.ny + j) * g.nz + k;
    else if (a == 1) o = ((size_t)i * w + s) * g.nz + k;
    else o = ((size_t)i * g.ny + j) * w + s;
    return (long long)(d.psl.off[slot] + o);
}

__device__ __forceinline__ float cls_cb(const MatDev &m, int c, size_t id, size_t ncell) {
    return m.mode == 0 ? m.lut[8 * (size_t)m.cls[id] + 2 * c + 1] : m.cb[(size_t)c * ncell + id];
}
__device__ __forceinline__ float cls_ca(const MatDev &m, int c, size_t id, size_t ncell) {
    return m.mode == 0 ? m.lut[8 * (size_t)m.cls[id] + 2 * c] : m.ca[(size_t)c * ncell + id];
}

// E update's pec factor of component c at (i, j, k)
__device__ __forceinline__ float pec_of(const GridDev &g, int c, int i, int j, int k) {
    const float px = axv(g.ax[0], i, AX_PEC), py = axv(g.ax[1], j, AX_PEC), pz = axv(g.ax[2], k, AX_PEC);
    return c == 0 ? py * pz : (c == 1 ? pz * px : px * py);
}

// q of an E-transpose term at cell (i, j, k): (1/kappa + a) sgn nu_c + a lpe, nu_c = cb pec lE_c
__device__ __forceinline__ float q_e(const AdjDev &d, const Term &t, int i, int j, int k, size_t id) {
    const AxisDev &A = d.g.ax[t.a];
    const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
    const float nu = cls_cb(d.mat, t.c, id, d.ncell) * pec_of(d.g, t.c, i, j, k) *
                     d.lE[(size_t)t.c * d.ncell + id];
    const long long pi = psi_at(d, t.slot, i, j, k);
    if (pi < 0) return (float)t.sgn * nu;
    const float ae = axv(A, p, AX_AE), ke = axv(A, p, AX_KE);
    return (ke + ae) * (float)t.sgn * nu + ae * d.lpe[pi];
}

// q of an H-transpose term: (1/kappa + a) f + a lph, f = -ch sgn lH_c
__device__ __forceinline__ float q_h(const AdjDev &d, const Term &t, int i, int j, int k, size_t id) {
    const AxisDev &A = d.g.ax[t.a];
    const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
    const float f = -d.ch * (float)t.sgn * d.lH[(size_t)t.c * d.ncell + id];
    const long long pi = psi_at(d, t.slot, i, j, k);
    if (pi < 0) return f;
    const float ah = axv(A, p, AX_AH), kh = axv(A, p, AX_KH);
    return (kh + ah) * f + ah * d.lph[pi];
}

// E transpose, part 1: lH_b += D_bwd^T q_{c,a} for the two terms with source b.
// D_bwd f[p] = (f[p] - f[prv p] mpv[p]) idd[p]  =>  (D_bwd^T q)[p] = q[p] idd[p]
//   - sum over p' with prv[p'] == p of q[p'] mpv[p'] idd[p']  (p' = nxt p, or p itself at a
//   clamped end)
__global__ void __launch_bounds__(OEM_ADJ_THREADS) k_adj_e1(AdjDev d)
{
    const int ny = d.g.ny, nz = d.g.nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < d.ncell;
         id += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(id, ny, nz, i, j, k);
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const Term t = kTerms[q];
            const AxisDev &A = d.g.ax[t.a];
            const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
            const float qs = q_e(d, t, i, j, k, id);
            float v = qs * axv(A, p, AX_IDLD);
            if (A.prv[p] == p) v -= qs * axv(A, p, AX_MPV) * axv(A, p, AX_IDLD);   // clamped end
            const int pn = A.nxt[p];
            if (pn != p && A.prv[pn] == p) {
                const int ii = t.a == 0 ? pn : i, jj = t.a == 1 ? pn : j, kk = t.a == 2 ? pn : k;
                const size_t idn = shift(id, t.a, p, pn, ny, nz);
                v -= q_e(d, t, ii, jj, kk, idn) * axv(A, pn, AX_MPV) * axv(A, pn, AX_IDLD);
            }
            acc[t.b] += v;
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) d.lH[(size_t)b * d.ncell + id] += acc[b];
    }
}

// E transpose, part 2 (cell-local): lpe <- b (sgn nu + lpe), lE <- ca pec lE
__global__ void __launch_bounds__(OEM_ADJ_THREADS) k_adj_e2(AdjDev d)
{
    const int ny = d.g.ny, nz = d.g.nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < d.ncell;
         id += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(id, ny, nz, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const Term t = kTerms[q];
            const long long pi = psi_at(d, t.slot, i, j, k);
            if (pi < 0) continue;
            const AxisDev &A = d.g.ax[t.a];
            const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
            const float nu = cls_cb(d.mat, t.c, id, d.ncell) * pec_of(d.g, t.c, i, j, k) *
                             d.lE[(size_t)t.c * d.ncell + id];
            d.lpe[pi] = axv(A, p, AX_BE) * ((float)t.sgn * nu + d.lpe[pi]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float &l = d.lE[(size_t)c * d.ncell + id];
            l = cls_ca(d.mat, c, id, d.ncell) * pec_of(d.g, c, i, j, k) * l;
        }
    }
}

// H transpose, part 1: lE_b += D_fwd^T q_{c,a}.
// D_fwd f[p] = (f[nxt p] mnx[p] - f[p]) idp[p]  =>  (D_fwd^T q)[p] = -q[p] idp[p]
//   + sum over p' with nxt[p'] == p of q[p'] mnx[p'] idp[p']  (p' = prv p, or p itself at a
//   clamped end)
__global__ void __launch_bounds__(OEM_ADJ_THREADS) k_adj_h1(AdjDev d)
{
    const int ny = d.g.ny, nz = d.g.nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < d.ncell;
         id += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(id, ny, nz, i, j, k);
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const Term t = kTerms[q];
            const AxisDev &A = d.g.ax[t.a];
            const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
            const float qs = q_h(d, t, i, j, k, id);
            float v = -qs * axv(A, p, AX_IDLP);
            if (A.nxt[p] == p) v += qs * axv(A, p, AX_MNX) * axv(A, p, AX_IDLP);   // clamped end
            const int pp = A.prv[p];
            if (pp != p && A.nxt[pp] == p) {
                const int ii = t.a == 0 ? pp : i, jj = t.a == 1 ? pp : j, kk = t.a == 2 ? pp : k;
                const size_t idp = shift(id, t.a, p, pp, ny, nz);
                v += q_h(d, t, ii, jj, kk, idp) * axv(A, pp, AX_MNX) * axv(A, pp, AX_IDLP);
            }
            acc[t.b] += v;
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) d.lE[(size_t)b * d.ncell + id] += acc[b];
    }
}

// H transpose, part 2 (cell-local): lph <- b (f + lph)
__global__ void __launch_bounds__(OEM_ADJ_THREADS) k_adj_h2(AdjDev d)
{
    const int ny = d.g.ny, nz = d.g.nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < d.ncell;
         id += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(id, ny, nz, i, j, k);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const Term t = kTerms[q];
            const long long pi = psi_at(d, t.slot, i, j, k);
            if (pi < 0) continue;
            const AxisDev &A = d.g.ax[t.a];
            const int p = t.a == 0 ? i : (t.a == 1 ? j : k);
            const float f = -d.ch * (float)t.sgn * d.lH[(size_t)t.c * d.ncell + id];
            d.lph[pi] = axv(A, p, AX_BH) * (f + d.lph[pi]);
        }
    }
}

// Objective: F[comp][cell] += Re(z * tab[n]) for one point list (scale 1).
__global__ void __launch_bounds__(256) k_adj_obj(float *F, PointsDev pd, MatDev mat, size_t ncell, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pd.n_groups) return;
    point_group<float>(F, pd, mat, 1.0f, ncell, n, t);
}

// eps gradient: g[q] += lE . (dca e0 + dcb (e1 - ca e0) / cb)
__global__ void __launch_bounds__(256)
k_adj_grad(const float *lE, const int *comp, const int *cell, const float *dca, const float *dcb,
           const float *ca, const float *cb, const float *e0, const float *e1, int n,
           size_t ncell, double *g)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float l = lE[(size_t)comp[q] * ncell + cell[q]];
    const float s = (e1[q] - ca[q] * e0[q]) / cb[q];
    g[q] += (double)l * ((double)dca[q] * e0[q] + (double)dcb[q] * s);
}

// rec[q] = E[comp[q]][cell[q]]
__global__ void __launch_bounds__(256)
k_adj_gather(const float *E, const int *comp, const int *cell, int n, size_t ncell, float *rec)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) rec[q] = E[(size_t)comp[q] * ncell + cell[q]];
}

}  // namespace

struct oem_adjoint_state {
    // checkpoints
    int n_slots = 0;
    float *ckpt = nullptr;           // n_slots * (6 ncell + 2 psi_total)
    std::vector<int64_t> ckpt_step;
    // adjoint fields
    float *lE = nullptr, *lH = nullptr, *lpe = nullptr, *lph = nullptr;
    // design
    int n_design = 0, max_seg = 0;
    int *d_comp = nullptr, *d_cell = nullptr;
    float *dca = nullptr, *dcb = nullptr, *ca = nullptr, *cb = nullptr;
    double *grad = nullptr;
    float *rec = nullptr;            // (max_seg + 1) * n_design
    int64_t rec_s0 = -1;
    int rec_n = 0;
    // objective point lists: E side, H side
    std::vector<Points> obj;
};

namespace oemi {
void adjoint_state_free(oem_adjoint_state *a) { delete a; }
}

static int adj_check(oem_ctx *ctx) {
    if (!ctx->prepared) return fail(OEM_ESTATE, "oem_prepare not called");
    if (ctx->cplx || ctx->has_disp)
        return fail(OEM_ENOTSUP, "the adjoint covers real, non-dispersive scenes");
    if (ctx->pool) return fail(OEM_ENOTSUP, "pooled contexts have no adjoint");
    if (!ctx->adj) ctx->adj = new oem_adjoint_state();
    return OEM_OK;
}

static size_t state_floats(oem_ctx *c) { return 6 * c->ncell + 2 * (size_t)c->psi.total; }

extern "C" int oem_checkpoint_alloc(oem_ctx *ctx, int32_t n_slots) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (n_slots < 1 || a->n_slots) return fail(OEM_EINVAL, "checkpoint slots %d (once per context)", n_slots);
    TRY(ctx->alloc(&a->ckpt, (size_t)n_slots * state_floats(ctx), false));
    a->n_slots = n_slots;
    a->ckpt_step.assign(n_slots, -1);
    return OEM_OK;
}

static int copy_state(oem_ctx *c, float *dst_base, bool save) {
    const size_t nc = c->ncell, pt = c->psi.total;
    const bool odd = c->fused && (c->step & 1);
    float *E = odd ? c->E2 : c->E, *H = odd ? c->H2 : c->H;
    const PsiDev &ps = odd ? c->psi2 : c->psi;
    float *parts[4] = {E, H, static_cast<float *>(ps.h), static_cast<float *>(ps.e)};
    const size_t n[4] = {3 * nc, 3 * nc, pt, pt};
    size_t o = 0;
    for (int q = 0; q < 4; ++q) {
        void *dst = save ? (void *)(dst_base + o) : (void *)parts[q];
        const void *src = save ? (const void *)parts[q] : (const void *)(dst_base + o);
        HIP_OK(hipMemcpyAsync(dst, src, n[q] * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        o += n[q];
    }
    return OEM_OK;
}

extern "C" int oem_checkpoint_save(oem_ctx *ctx, int32_t slot) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (slot < 0 || slot >= a->n_slots) return fail(OEM_EINVAL, "checkpoint slot %d", slot);
    if (ctx->pending) return fail(OEM_ESTATE, "checkpoint between DFT batches only");
    TRY(copy_state(ctx, a->ckpt + (size_t)slot * state_floats(ctx), true));
    a->ckpt_step[slot] = ctx->step;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_checkpoint_load(oem_ctx *ctx, int32_t slot) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (slot < 0 || slot >= a->n_slots || a->ckpt_step[slot] < 0)
        return fail(OEM_EINVAL, "checkpoint slot %d is empty", slot);
    ctx->step = a->ckpt_step[slot];
    TRY(copy_state(ctx, a->ckpt + (size_t)slot * state_floats(ctx), false));
    const int st[2] = {(int)ctx->step, (int)ctx->step};
    HIP_OK(hipMemcpyAsync(ctx->d_step, st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    ctx->pending = 0;
    return OEM_OK;
}

extern "C" int oem_adjoint_setup(oem_ctx *ctx, int32_t n_design, const int32_t *comp,
                                 const int32_t *cell, const float *dca, const float *dcb,
                                 const float *ca, const float *cb, int32_t max_seg) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (a->lE) return fail(OEM_ESTATE, "adjoint already set up");
    if (n_design < 1 || max_seg < 1 || !comp || !cell || !dca || !dcb || !ca || !cb)
        return fail(OEM_EINVAL, "bad design tables");
    for (int q = 0; q < n_design; ++q) {
        if (comp[q] < 0 || comp[q] > 2 || cell[q] < 0 || (size_t)cell[q] >= ctx->ncell)
            return fail(OEM_EINVAL, "design entry %d out of range", q);
        if (cb[q] == 0.f) return fail(OEM_EINVAL, "design entry %d has cb = 0 (PEC)", q);
    }
    const size_t nc = ctx->ncell;
    TRY(ctx->alloc(&a->lE, 3 * nc, true));
    TRY(ctx->alloc(&a->lH, 3 * nc, true));
    TRY(ctx->alloc(&a->lpe, (size_t)ctx->psi.total, true));
    TRY(ctx->alloc(&a->lph, (size_t)ctx->psi.total, true));
    TRY(ctx->upload(&a->d_comp, (const int *)comp, (size_t)n_design));
    TRY(ctx->upload(&a->d_cell, (const int *)cell, (size_t)n_design));
    TRY(ctx->upload(&a->dca, dca, (size_t)n_design));
    TRY(ctx->upload(&a->dcb, dcb, (size_t)n_design));
    TRY(ctx->upload(&a->ca, ca, (size_t)n_design));
    TRY(ctx->upload(&a->cb, cb, (size_t)n_design));
    TRY(ctx->alloc(&a->grad, (size_t)n_design, true));
    TRY(ctx->alloc(&a->rec, (size_t)(max_seg + 1) * n_design, false));
    a->n_design = n_design;
    a->max_seg = max_seg;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_adjoint_objective(oem_ctx *ctx, int32_t magnetic, int32_t n_pts,
                                     const int32_t *comp, const int32_t *cell, const float *z,
                                     const double *tab) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    if (n_pts < 1 || !comp || !cell || !z || !tab) return fail(OEM_EINVAL, "bad objective tables");
    for (int q = 0; q < n_pts; ++q)
        if (comp[q] < 0 || comp[q] > 2 || cell[q] < 0 || (size_t)cell[q] >= ctx->ncell)
            return fail(OEM_EINVAL, "objective point %d out of range", q);
    std::vector<int32_t> col(n_pts, 0);
    std::vector<float> one(n_pts, 1.f);
    Points P{};
    TRY(oemi::build_points(ctx, magnetic, PT_TFSF, 1, 1, n_pts, comp, cell, one.data(), col.data(),
                           z, 1, tab, (size_t)ctx->dims.n_steps * 2, &P));
    P.d.e_side = 0;                      // scale given by the kernel argument (1)
    ctx->adj->obj.push_back(P);
    return OEM_OK;
}

extern "C" int oem_adjoint_record(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (!a->rec) return fail(OEM_ESTATE, "oem_adjoint_setup not called");
    if (n_steps < 1 || n_steps > a->max_seg) return fail(OEM_EINVAL, "segment of %lld steps", (long long)n_steps);
    if (ctx->step + n_steps > ctx->dims.n_steps) return fail(OEM_EINVAL, "beyond the step tables");
    const dim3 b(256), g((a->n_design + 255) / 256);
    ctx->mons_off = true;
    a->rec_s0 = ctx->step;
    hipLaunchKernelGGL(k_adj_gather, g, b, 0, ctx->stream, ctx->cur_E(), a->d_comp, a->d_cell,
                       a->n_design, ctx->ncell, a->rec);
    int rc = OEM_OK;
    for (int64_t s = 0; s < n_steps && !rc; ++s) {
        rc = oemi::launch_one_step(ctx, false);
        if (!rc)
            hipLaunchKernelGGL(k_adj_gather, g, b, 0, ctx->stream, ctx->cur_E(), a->d_comp,
                               a->d_cell, a->n_design, ctx->ncell, a->rec + (size_t)(s + 1) * a->n_design);
    }
    ctx->mons_off = false;
    ctx->pre_ready = false;
    TRY(rc);
    HIP_OK(hipGetLastError());
    a->rec_n = (int)n_steps;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_adjoint_back(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (a->rec_s0 < 0 || n_steps != a->rec_n)
        return fail(OEM_ESTATE, "back over exactly the recorded segment (%d steps)", a->rec_n);
    AdjDev d{};
    d.lE = a->lE; d.lH = a->lH; d.lpe = a->lpe; d.lph = a->lph;
    d.psl = ctx->psi;
    d.g = ctx->g;
    d.mat = ctx->mat;
    d.ch = ctx->dims.ch;
    d.ncell = ctx->ncell;
    const dim3 tb(OEM_ADJ_THREADS);
    const dim3 tg((unsigned)std::min<size_t>(OEM_ADJ_BLOCKS, (ctx->ncell + OEM_ADJ_THREADS - 1) / OEM_ADJ_THREADS));
    const dim3 gb(256), gg((a->n_design + 255) / 256);
    for (int s = a->rec_n - 1; s >= 0; --s) {
        const int n = (int)(a->rec_s0 + s);          // forward step n: E^n -> E^{n+1}
        for (auto &P : a->obj)
            hipLaunchKernelGGL(k_adj_obj, dim3((P.d.n_groups + 255) / 256), dim3(256), 0,
                               ctx->stream, P.magnetic ? a->lH : a->lE, P.d, ctx->mat,
                               ctx->ncell, n);
        hipLaunchKernelGGL(k_adj_grad, gg, gb, 0, ctx->stream, a->lE, a->d_comp, a->d_cell,
                           a->dca, a->dcb, a->ca, a->cb, a->rec + (size_t)s * a->n_design,
                           a->rec + (size_t)(s + 1) * a->n_design, a->n_design, ctx->ncell,
                           a->grad);
        hipLaunchKernelGGL(k_adj_e1, tg, tb, 0, ctx->stream, d);
        hipLaunchKernelGGL(k_adj_e2, tg, tb, 0, ctx->stream, d);
        hipLaunchKernelGGL(k_adj_h1, tg, tb, 0, ctx->stream, d);
        hipLaunchKernelGGL(k_adj_h2, tg, tb, 0, ctx->stream, d);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(ctx->stream));
    a->rec_s0 = -1;
    a->rec_n = 0;
    return OEM_OK;
}

extern "C" int oem_adjoint_gradient(oem_ctx *ctx, double *out, size_t n) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (!a->grad || !out || n != (size_t)a->n_design)
        return fail(OEM_EINVAL, "gradient holds %d values", a->n_design);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipMemcpy(out, a->grad, n * sizeof(double), hipMemcpyDeviceToHost));
    return OEM_OK;
}
