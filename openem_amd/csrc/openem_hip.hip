// C-ABI implementation of include/openem_hip.h for MI355X (gfx950).
//
// One context = one scene resident in HBM: fields, CPML psi slabs, material classes, ADE pole
// state, source and monitor tables.  oem_run advances it with the launch sequence of the
// reference's _step_body (`openem/solver.py:1145-1162`):
//
//   update_h (+ plane-wave H correction) -> magnetic dipoles / TF/SF H faces -> ADE pre ->
//   update_e (+ plane-wave E correction) -> TF/SF E faces -> electric dipoles -> ADE post
//   (fused with the next pre) -> DFT snapshots -> step++
//
// with the DFT flushed every dft_batch steps.  Outside a serving pool, update_h and update_e run
// as one fused kernel (update_he) that ping-pongs between two field sets, the H-side point
// corrections added to the old H just before it.  Blocks of dft_batch steps aligned to the batch
// are captured once into a hipGraph and replayed; the device-side step counter lets one graph
// serve every block (the reference does the same with a CUDA graph, `openem/solver.py:1165-1182`).

#include "oem_internal.h"

#include <array>

__global__ void k_step_advance(int *stp) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *stp += 1;
}

std::mutex &oemi::device_sync_mutex() {
    static std::mutex m;
    return m;
}

extern "C" int oem_version(void) { return OEM_ABI_VERSION; }

extern "C" int oem_last_error(char *buf, size_t n) {
    if (!buf || n == 0) return OEM_EINVAL;
    std::snprintf(buf, n, "%s", oemi::g_err.c_str());
    return OEM_OK;
}

extern "C" int oem_device_count(int32_t *n) {
    if (!n) return fail(OEM_EINVAL, "null output");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return OEM_OK;
}

int ctx_init(oem_ctx *c, const oem_dims *d, int32_t device) {
    if (d->nx < 1 || d->ny < 1 || d->nz < 1)
        return fail(OEM_EINVAL, "grid %d x %d x %d", d->nx, d->ny, d->nz);
    if ((int64_t)d->ny * d->nz > (int64_t)1 << 30)
        return fail(OEM_EINVAL, "plane ny*nz = %lld too large", (long long)d->ny * d->nz);
    if (d->n_steps < 1) return fail(OEM_EINVAL, "n_steps %d", d->n_steps);
    if (d->dft_batch < 1 || d->dft_batch > 64)
        return fail(OEM_EINVAL, "dft_batch %d outside 1..64", d->dft_batch);
    const int n[3] = {d->nx, d->ny, d->nz};
    for (int a = 0; a < 3; ++a)
        if (d->pml_lo[a] < 0 || d->pml_hi[a] < 0 || d->pml_lo[a] + d->pml_hi[a] > n[a])
            return fail(OEM_EINVAL, "axis %d: pml %d+%d exceeds %d cells", a, d->pml_lo[a],
                        d->pml_hi[a], n[a]);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(OEM_EHIP, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(OEM_EINVAL, "device %d of %d", device, ndev);
    HIP_OK(hipSetDevice(device));
    c->device = device;
    c->dims = *d;
    c->ncell = (size_t)d->nx * d->ny * d->nz;
    c->cplx = d->complex_fields != 0;
    c->esz = c->cplx ? 2 : 1;
    {
        std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            return fail(OEM_EHIP, "hipStreamCreate failed");
    }
    const size_t nc = c->ncell;
    if (nc * 4 * c->esz > 0xffffffffull)
        return fail(OEM_ENOTSUP, "%zu cells: one field component exceeds 4 GiB (limit 2^%d cells)",
                    nc, c->cplx ? 29 : 30);
    const size_t slab[3] = {(size_t)(d->pml_lo[0] + d->pml_hi[0]) * d->ny * d->nz,
                            (size_t)d->nx * (d->pml_lo[1] + d->pml_hi[1]) * d->nz,
                            (size_t)d->nx * d->ny * (d->pml_lo[2] + d->pml_hi[2])};
    size_t total = 0;
    for (int a = 0; a < 3; ++a)
        for (int m = 0; m < 2; ++m) {
            c->psi.off[2 * a + m] = (uint32_t)total;
            total += slab[a];
        }
    if (total * 4 * c->esz > 0xffffffffull) return fail(OEM_ENOTSUP, "CPML slabs exceed 4 GiB per side");
    c->psi.total = (uint32_t)(total ? total : 1);
    // E and H in one allocation (H right after E), and the H- and E-side psi in one: a kernel
    // then reaches a whole field set, or a whole psi set, through one buffer descriptor
    TRY(c->alloc(&c->E, 6 * nc * c->esz, true));
    c->H = c->E + 3 * nc * c->esz;
    float *pp = nullptr;
    TRY(c->alloc(&pp, 2 * (size_t)c->psi.total * c->esz, true));
    c->psi.h = pp;
    c->psi.e = pp + (size_t)c->psi.total * c->esz;
    TRY(c->alloc(&c->d_step, 2, true));
    HIP_OK(hipStreamSynchronize(c->stream));
    return OEM_OK;
}

extern "C" int oem_ctx_create(const oem_dims *d, int32_t device, oem_ctx **out) {
    if (!d || !out) return fail(OEM_EINVAL, "null argument");
    *out = nullptr;
    oem_ctx *c = new oem_ctx();
    const int rc = ctx_init(c, d, device);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return OEM_OK;
}

extern "C" int oem_ctx_destroy(oem_ctx *ctx) {
    if (!ctx) return OEM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->pool) return oemi::pool_retire(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return OEM_OK;
}

extern "C" int oem_set_chunk(oem_ctx *ctx, int32_t chunk) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (chunk < 0) return fail(OEM_EINVAL, "chunk %d", chunk);
    ctx->chunk_req = chunk;
    return OEM_OK;
}

extern "C" int oem_set_step_form(oem_ctx *ctx, int32_t form) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (form < 0 || form > 2) return fail(OEM_EINVAL, "step form %d", form);
    if (form == 2 && ctx->pool) return fail(OEM_ENOTSUP, "a pooled context steps with the two sweeps");
    ctx->form_req = form;
    return OEM_OK;
}

extern "C" int oem_launch_plan(oem_ctx *ctx, int32_t *form, int32_t *chunk) {
    CHECK_CTX(ctx);
    if (!ctx->prepared) return fail(OEM_ESTATE, "oem_prepare not called");
    if (!form || !chunk) return fail(OEM_EINVAL, "null output");
    *form = ctx->fused ? 2 : 1;
    *chunk = ctx->fused ? ctx->fg.chunk : ctx->g.chunk;
    return OEM_OK;
}

extern "C" int oem_set_axis(oem_ctx *ctx, int32_t axis, const float *idl_p, const float *idl_d,
                            const int32_t *nxt, const int32_t *prv, const float *mnx,
                            const float *mpv, const float *pec, const float *a_e,
                            const float *b_e, const float *k_e, const float *a_h,
                            const float *b_h, const float *k_h, const float *mnx_im,
                            const float *mpv_im) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (axis < 0 || axis > 2) return fail(OEM_EINVAL, "axis %d", axis);
    const int n = axis == 0 ? ctx->dims.nx : (axis == 1 ? ctx->dims.ny : ctx->dims.nz);
    const float *fa[] = {idl_p, idl_d, mnx, mpv, pec, a_e, b_e, k_e, a_h, b_h, k_h};
    for (const float *p : fa)
        if (!p) return fail(OEM_EINVAL, "axis %d: null table", axis);
    if (!nxt || !prv) return fail(OEM_EINVAL, "axis %d: null neighbour table", axis);
    for (int i = 0; i < n; ++i)
        if (nxt[i] < 0 || nxt[i] >= n || prv[i] < 0 || prv[i] >= n)
            return fail(OEM_EINVAL, "axis %d: neighbour index out of range at %d", axis, i);
    // the x march carries plane i+1 (H) / i-1 (E) between iterations: interior neighbours
    // must be the adjacent planes
    for (int i = 0; i + 1 < n; ++i)
        if (nxt[i] != i + 1 || prv[i + 1] != i)
            return fail(OEM_EINVAL, "axis %d: interior neighbours must be i+-1 (at %d)", axis, i);
    AxisDev &A = ctx->g.ax[axis];
    std::vector<float> tab((size_t)n * AX_W, 0.f);
    const int fid[11] = {AX_IDLP, AX_IDLD, AX_MNX, AX_MPV, AX_PEC, AX_AE,
                         AX_BE,   AX_KE,   AX_AH,  AX_BH,  AX_KH};
    for (int t = 0; t < 11; ++t)
        for (int i = 0; i < n; ++i) tab[(size_t)i * AX_W + fid[t]] = fa[t][i];
    for (int i = 0; i < n; ++i) {
        const float a = mnx_im ? mnx_im[i] : 0.f, b = mpv_im ? mpv_im[i] : 0.f;
        if ((a != 0.f || b != 0.f) && !ctx->cplx)
            return fail(OEM_EINVAL, "axis %d: complex (Bloch) masks need complex fields", axis);
        tab[(size_t)i * AX_W + AX_MNXI] = a;
        tab[(size_t)i * AX_W + AX_MPVI] = b;
    }
    float *dt = nullptr;
    TRY(ctx->upload(&dt, tab.data(), tab.size()));
    A.tab = dt;
    int *pn = nullptr, *pp = nullptr;
    TRY(ctx->upload(&pn, (const int *)nxt, (size_t)n));
    TRY(ctx->upload(&pp, (const int *)prv, (size_t)n));
    A.nxt = pn;
    A.prv = pp;
    A.n = n;
    A.lo = ctx->dims.pml_lo[axis];
    A.hi = ctx->dims.pml_hi[axis];
    ctx->axis_set[axis] = true;
    return OEM_OK;
}

extern "C" int oem_set_materials_lut(oem_ctx *ctx, const uint16_t *cls, int32_t n_classes,
                                     const float *lut) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (!cls || !lut || n_classes < 1 || n_classes > 65535)
        return fail(OEM_EINVAL, "bad class table (n_classes=%d)", n_classes);
    for (size_t i = 0; i < ctx->ncell; ++i)
        if (cls[i] >= n_classes) return fail(OEM_EINVAL, "class %u at cell %zu >= %d", cls[i], i, n_classes);
    uint16_t *dc = nullptr;
    float *dl = nullptr;
    TRY(ctx->upload(&dc, cls, ctx->ncell));
    TRY(ctx->upload(&dl, lut, (size_t)n_classes * 8));
    ctx->class_flagged.assign(n_classes, 0);
    for (int q = 0; q < n_classes; ++q) {
        uint32_t fl;
        std::memcpy(&fl, lut + 8 * (size_t)q + 6, 4);
        if (fl & ~7u) return fail(OEM_EINVAL, "class %d: flag bits 0x%x beyond 0..2", q, fl);
        ctx->class_flagged[q] = fl != 0;
    }
    ctx->mat.mode = 0;
    ctx->mat.n_classes = n_classes;
    ctx->mat.cls = dc;
    ctx->mat.lut = dl;
    ctx->mat_set = true;
    return OEM_OK;
}

extern "C" int oem_set_materials_dense(oem_ctx *ctx, const float *ca, const float *cb) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (!ca || !cb) return fail(OEM_EINVAL, "null ca/cb");
    float *a = nullptr, *b = nullptr;
    TRY(ctx->upload(&a, ca, 3 * ctx->ncell));
    TRY(ctx->upload(&b, cb, 3 * ctx->ncell));
    ctx->mat.mode = 1;
    ctx->mat.ca = a;
    ctx->mat.cb = b;
    ctx->mat_set = true;
    return OEM_OK;
}

extern "C" int oem_set_dispersion(oem_ctx *ctx, int32_t n_entry, const int32_t *comp,
                                  const int32_t *cell, const int32_t *ofs, int32_t n_pole,
                                  const int32_t *cidx, int32_t n_coef, const float *am1_re,
                                  const float *am1_im, const float *b_re, const float *b_im,
                                  const int32_t *hist_box) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (n_entry < 1 || n_pole < 0 || n_coef < 1 || !comp || !cell || !ofs || !hist_box ||
        (n_pole && !cidx) || !am1_re || !am1_im || !b_re || !b_im)
        return fail(OEM_EINVAL, "bad dispersion tables");
    const oem_dims &d = ctx->dims;
    const int hb[6] = {hist_box[0], hist_box[1], hist_box[2], hist_box[3], hist_box[4], hist_box[5]};
    if (hb[0] < 0 || hb[1] > d.nx || hb[0] >= hb[1] || hb[2] < 0 || hb[3] > d.ny ||
        hb[2] >= hb[3] || hb[4] < 0 || hb[5] > d.nz || hb[4] >= hb[5])
        return fail(OEM_EINVAL, "hist box outside the grid");
    if (ofs[0] != 0 || ofs[n_entry] != n_pole) return fail(OEM_EINVAL, "ofs endpoints");
    const int pl = d.ny * d.nz;
    for (int e = 0; e < n_entry; ++e) {
        if (ofs[e + 1] < ofs[e]) return fail(OEM_EINVAL, "ofs not monotonic at %d", e);
        if (comp[e] < 0 || comp[e] > 2) return fail(OEM_EINVAL, "comp out of range at %d", e);
        if (cell[e] < 0 || (size_t)cell[e] >= ctx->ncell)
            return fail(OEM_EINVAL, "cell out of range at %d", e);
        const int i = cell[e] / pl, r = cell[e] % pl, j = r / d.nz, k = r % d.nz;
        if (i < hb[0] || i >= hb[1] || j < hb[2] || j >= hb[3] || k < hb[4] || k >= hb[5])
            return fail(OEM_EINVAL, "entry %d outside the hist box", e);
    }
    for (int t = 0; t < n_pole; ++t)
        if (cidx[t] < 0 || cidx[t] >= n_coef) return fail(OEM_EINVAL, "cidx out of range at %d", t);
    const size_t hsz = 3ull * (hb[1] - hb[0]) * (hb[3] - hb[2]) * (hb[5] - hb[4]);
    if (hsz * 4 * ctx->esz > 0xffffffffull) return fail(OEM_ENOTSUP, "ADE history box exceeds 4 GiB");
    DispDev &D = ctx->disp;
    D = DispDev{};
    D.n_entry = n_entry;
    // the compact form when every entry has the same pole count (1..4)
    const int np0 = ofs[1] - ofs[0];
    const char *csr = std::getenv("OEM_ADE_CSR");   // "1": keep the CSR form
    bool uniform = np0 >= 1 && np0 <= 4 && ctx->ncell <= (1ull << 30) && n_coef <= 65536 &&
                   !(csr && csr[0] == '1');
    for (int e = 1; uniform && e < n_entry; ++e) uniform = ofs[e + 1] - ofs[e] == np0;
    if (uniform) {
        const int bx = hb[1] - hb[0], by = hb[3] - hb[2], bz = hb[5] - hb[4];
        std::vector<uint32_t> meta(2 * (size_t)n_entry);
        for (int e = 0; e < n_entry; ++e) {
            const int i = cell[e] / pl, r = cell[e] % pl, j = r / d.nz, k = r % d.nz;
            meta[2 * e] = (uint32_t)comp[e] << 30 | (uint32_t)cell[e];
            meta[2 * e + 1] = (uint32_t)((((size_t)comp[e] * bx + (i - hb[0])) * by + (j - hb[2])) * bz +
                                         (k - hb[4]));
        }
        std::vector<uint16_t> c16((size_t)n_pole);
        for (int t = 0; t < n_pole; ++t) c16[t] = (uint16_t)cidx[t];
        std::vector<float> coef(4 * (size_t)n_coef);
        for (int q = 0; q < n_coef; ++q) {
            coef[4 * q] = am1_re[q]; coef[4 * q + 1] = am1_im[q];
            coef[4 * q + 2] = b_re[q]; coef[4 * q + 3] = b_im[q];
        }
        uint32_t *p_meta;
        uint16_t *p_c16;
        float *p_coef;
        TRY(ctx->upload(&p_meta, meta.data(), meta.size()));
        TRY(ctx->upload(&p_c16, c16.data(), c16.size()));
        TRY(ctx->upload(&p_coef, coef.data(), coef.size()));
        D.np = np0;
        D.meta = reinterpret_cast<const uint2 *>(p_meta);
        D.c16 = p_c16;
        D.coef = reinterpret_cast<const float4 *>(p_coef);
        TRY(ctx->alloc(&D.sp, (size_t)n_pole, true));
        if (ctx->cplx) TRY(ctx->alloc(&D.sq, (size_t)n_pole, true));
    } else {
    int *p_comp, *p_cell, *p_ofs, *p_cidx;
    float *ar, *ai, *br, *bi;
    TRY(ctx->upload(&p_comp, (const int *)comp, (size_t)n_entry));
    TRY(ctx->upload(&p_cell, (const int *)cell, (size_t)n_entry));
    TRY(ctx->upload(&p_ofs, (const int *)ofs, (size_t)n_entry + 1));
    if (n_pole) {
        TRY(ctx->upload(&p_cidx, (const int *)cidx, (size_t)n_pole));
    } else {
        TRY(ctx->alloc(&p_cidx, 1, true));
    }
    TRY(ctx->upload(&ar, am1_re, (size_t)n_coef));
    TRY(ctx->upload(&ai, am1_im, (size_t)n_coef));
    TRY(ctx->upload(&br, b_re, (size_t)n_coef));
    TRY(ctx->upload(&bi, b_im, (size_t)n_coef));
    D.comp = p_comp; D.cell = p_cell; D.ofs = p_ofs; D.cidx = p_cidx;
    D.am1_re = ar; D.am1_im = ai; D.b_re = br; D.b_im = bi;
    TRY(ctx->alloc(&D.p_re, (size_t)n_pole, true));
    TRY(ctx->alloc(&D.p_im, (size_t)n_pole, true));
    if (ctx->cplx) {
        TRY(ctx->alloc(&D.q_re, (size_t)n_pole, true));
        TRY(ctx->alloc(&D.q_im, (size_t)n_pole, true));
    }
    }
    float *hist = nullptr;
    TRY(ctx->alloc(&hist, hsz * ctx->esz, true));
    D.hist = hist;
    for (int q = 0; q < 6; ++q) { D.hb[q] = hb[q]; ctx->mat.hb[q] = hb[q]; }
    ctx->mat.hist = D.hist;
    ctx->has_disp = true;
    ctx->disp_n_pole = n_pole;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_add_plane_source(oem_ctx *ctx, int32_t pol, int32_t kh, int32_t ks,
                                    float coef_h, float coef_e, const double *ex_inc,
                                    const double *hy_inc) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (ctx->pw.n >= OEM_MAX_PW) return fail(OEM_ENOTSUP, "at most %d plane sources", OEM_MAX_PW);
    if (pol < 0 || pol > 1) return fail(OEM_ENOTSUP, "plane wave polarisation %d (Ex/Ey only)", pol);
    if (kh < 0 || kh >= ctx->dims.nz || ks < 0 || ks >= ctx->dims.nz)
        return fail(OEM_EINVAL, "plane indices kh=%d ks=%d outside nz=%d", kh, ks, ctx->dims.nz);
    if (!ex_inc || !hy_inc) return fail(OEM_EINVAL, "null incident table");
    double *ex, *hy;
    TRY(ctx->upload(&ex, ex_inc, (size_t)ctx->dims.n_steps + 1));
    TRY(ctx->upload(&hy, hy_inc, (size_t)ctx->dims.n_steps + 1));
    const int s = ctx->pw.n++;
    ctx->pw.pol[s] = pol;
    ctx->pw.kh[s] = kh;
    ctx->pw.ks[s] = ks;
    ctx->pw.coef_h[s] = coef_h;
    ctx->pw.coef_e[s] = coef_e;
    ctx->pw.ex_inc[s] = ex;
    ctx->pw.hy_inc[s] = hy;
    return OEM_OK;
}

// Group points by (comp, cell), keeping input order inside a group, and upload them.
int oemi::build_points(oem_ctx *ctx, int magnetic, int kind, int tfsf, int cplx_amp,
                       int32_t n_pts, const int32_t *comp, const int32_t *cell, const float *coef,
                       const int32_t *col, const float *phase, int32_t ncol, const double *tab,
                       size_t tab_doubles, Points *out) {
    std::vector<int> order(n_pts);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return comp[a] != comp[b] ? comp[a] < comp[b] : cell[a] < cell[b];
    });
    std::vector<int> gc, gl, go, cl;
    std::vector<float> cf, ph;
    for (int t = 0; t < n_pts; ++t) {
        const int q = order[t];
        if (t == 0 || comp[q] != gc.back() || cell[q] != gl.back()) {
            gc.push_back(comp[q]);
            gl.push_back(cell[q]);
            go.push_back(t);
        }
        cf.push_back(coef[q]);
        cl.push_back(col[q]);
        if (phase) {
            ph.push_back(phase[2 * q]);
            ph.push_back(phase[2 * q + 1]);
        }
    }
    go.push_back(n_pts);
    Points P{};
    P.magnetic = magnetic ? 1 : 0;
    P.kind = kind;
    PointsDev &d = P.d;
    d.n_groups = (int)gc.size();
    d.nsteps1 = ctx->dims.n_steps + 1;
    d.ncol = ncol;
    d.e_side = magnetic ? 0 : 1;
    d.tfsf = tfsf;
    d.cplx_amp = cplx_amp;
    int *pa, *pb, *po, *pc;
    float *pf, *pp = nullptr;
    double *pt;
    TRY(ctx->upload(&pa, gc.data(), gc.size()));
    TRY(ctx->upload(&pb, gl.data(), gl.size()));
    TRY(ctx->upload(&po, go.data(), go.size()));
    TRY(ctx->upload(&pf, cf.data(), cf.size()));
    TRY(ctx->upload(&pc, cl.data(), cl.size()));
    if (phase) TRY(ctx->upload(&pp, ph.data(), ph.size()));
    TRY(ctx->upload(&pt, tab, tab_doubles));
    d.g_comp = pa; d.g_cell = pb; d.g_ofs = po; d.coef = pf; d.col = pc; d.phase = pp; d.tab = pt;
    P.cells = gl;
    *out = P;
    return OEM_OK;
}

static int add_points(oem_ctx *ctx, int magnetic, int kind, int tfsf, int cplx_amp, int32_t n_pts,
                      const int32_t *comp, const int32_t *cell, const float *coef,
                      const int32_t *col, const float *phase, int32_t ncol, const double *tab,
                      size_t tab_doubles) {
    Points P{};
    TRY(oemi::build_points(ctx, magnetic, kind, tfsf, cplx_amp, n_pts, comp, cell, coef, col,
                           phase, ncol, tab, tab_doubles, &P));
    ctx->points.push_back(P);
    return OEM_OK;
}

static int check_points(oem_ctx *ctx, int32_t n_pts, const int32_t *comp, const int32_t *cell,
                        const float *coef, const int32_t *col, int32_t ncol) {
    if (n_pts < 1 || !comp || !cell || !coef || !col) return fail(OEM_EINVAL, "bad point tables");
    for (int q = 0; q < n_pts; ++q) {
        if (comp[q] < 0 || comp[q] > 2) return fail(OEM_EINVAL, "point %d: component %d", q, comp[q]);
        if (cell[q] < 0 || (size_t)cell[q] >= ctx->ncell) return fail(OEM_EINVAL, "point %d: cell", q);
        if (col[q] < 0 || col[q] >= ncol) return fail(OEM_EINVAL, "point %d: column %d of %d", q, col[q], ncol);
    }
    return OEM_OK;
}

extern "C" int oem_add_dipoles(oem_ctx *ctx, int32_t magnetic, int32_t n_pts,
                               const int32_t *comp, const int32_t *cell, const float *coef,
                               const int32_t *src_of, int32_t n_src, const double *amp,
                               int32_t amp_complex) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (n_src < 1 || !amp) return fail(OEM_EINVAL, "bad dipole amplitude table");
    TRY(check_points(ctx, n_pts, comp, cell, coef, src_of, n_src));
    const size_t nd = (size_t)n_src * (ctx->dims.n_steps + 1) * (amp_complex ? 2 : 1);
    return add_points(ctx, magnetic, PT_DIPOLE, 0, amp_complex ? 1 : 0, n_pts, comp, cell, coef,
                      src_of, nullptr, n_src, amp, nd);
}

extern "C" int oem_add_tfsf_points(oem_ctx *ctx, int32_t magnetic, int32_t n_pts,
                                   const int32_t *comp, const int32_t *cell, const float *coef,
                                   const int32_t *col, const float *phase, int32_t ncol,
                                   const double *table) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (ncol < 1 || !table || !phase) return fail(OEM_EINVAL, "bad TF/SF table");
    TRY(check_points(ctx, n_pts, comp, cell, coef, col, ncol));
    if (!ctx->cplx)
        for (int q = 0; q < n_pts; ++q)
            if (phase[2 * q + 1] != 0.f)
                return fail(OEM_EINVAL, "point %d: complex phase needs complex fields", q);
    const size_t nd = (size_t)(ctx->dims.n_steps + 1) * ncol * 2;
    return add_points(ctx, magnetic, PT_TFSF, 1, 1, n_pts, comp, cell, coef, col, phase, ncol,
                      table, nd);
}

extern "C" int oem_add_mode_points(oem_ctx *ctx, int32_t magnetic, int32_t n_pts,
                                   const int32_t *comp, const int32_t *cell, const float *coef,
                                   const float *phase, const double *amp) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (!phase || !amp) return fail(OEM_EINVAL, "bad mode tables");
    std::vector<int32_t> col(n_pts > 0 ? n_pts : 0, 0);
    TRY(check_points(ctx, n_pts, comp, cell, coef, col.data(), 1));
    const size_t nd = (size_t)(ctx->dims.n_steps + 1) * 2;
    return add_points(ctx, magnetic, PT_MODE, 1, 1, n_pts, comp, cell, coef, col.data(), phase, 1,
                      amp, nd);
}

static int add_monitor(oem_ctx *ctx, MonDev m, const double *phase, int32_t *mon_id) {
    if (!phase) return fail(OEM_EINVAL, "null phase table");
    if (m.nf < 1 || m.nf > 65535) return fail(OEM_EINVAL, "nf=%d outside 1..65535", m.nf);
    m.T = ctx->dims.dft_batch;
    const size_t cnt = (size_t)m.ncomp * m.nf * m.npts;
    float *snap = nullptr;
    TRY(ctx->alloc(&snap, (size_t)m.T * m.ncomp * m.npts * ctx->esz, true));
    m.snap = snap;
    TRY(ctx->alloc(&m.re, cnt, true));
    TRY(ctx->alloc(&m.im, cnt, true));
    double *ph;
    TRY(ctx->upload(&ph, phase, (size_t)ctx->dims.n_steps * m.nf * 4));
    m.phase = ph;
    Monitor mon{m, cnt};
    ctx->mons.push_back(mon);
    if (mon_id) *mon_id = (int32_t)ctx->mons.size() - 1;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_add_flux_plane(oem_ctx *ctx, int32_t axis, int32_t km, int32_t nf,
                                  const double *phase, int32_t *mon_id) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    const int n[3] = {ctx->dims.nx, ctx->dims.ny, ctx->dims.nz};
    if (axis < 0 || axis > 2) return fail(OEM_EINVAL, "axis %d", axis);
    if (km < 1 || km >= n[axis]) return fail(OEM_EINVAL, "plane %d needs 1 <= km < %d", km, n[axis]);
    MonDev m{};
    m.kind = 0;
    m.axis = axis;
    m.km = km;
    m.n1 = axis == 0 ? n[1] : n[0];
    m.n2 = axis == 2 ? n[1] : n[2];
    m.ncomp = 4;
    m.npts = m.n1 * m.n2;
    m.nf = nf;
    return add_monitor(ctx, m, phase, mon_id);
}

extern "C" int oem_add_field_box(oem_ctx *ctx, const int32_t *origin, const int32_t *box,
                                 int32_t nf, const double *phase, int32_t *mon_id) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (!origin || !box) return fail(OEM_EINVAL, "null box");
    const int n[3] = {ctx->dims.nx, ctx->dims.ny, ctx->dims.nz};
    MonDev m{};
    m.kind = 1;
    size_t pts = 1;
    for (int a = 0; a < 3; ++a) {
        if (origin[a] < 0 || box[a] < 1 || origin[a] + box[a] > n[a])
            return fail(OEM_EINVAL, "box outside the grid on axis %d", a);
        m.o[a] = origin[a];
        m.b[a] = box[a];
        pts *= (size_t)box[a];
    }
    if (pts > (size_t)1 << 30) return fail(OEM_EINVAL, "box too large");
    m.ncomp = 6;
    m.npts = (int)pts;
    m.nf = nf;
    return add_monitor(ctx, m, phase, mon_id);
}

extern "C" int oem_add_time_monitor(oem_ctx *ctx, int32_t n_comps, const int32_t *comps,
                                    const int32_t *origin, const int32_t *box,
                                    int32_t step_begin, int32_t step_end, int32_t interval,
                                    int32_t n_slots, int32_t *mon_id) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (!comps || !origin || !box || !mon_id) return fail(OEM_EINVAL, "null argument");
    if (n_comps < 1 || n_comps > 6) return fail(OEM_EINVAL, "%d components", n_comps);
    if (interval < 1 || n_slots < 1 || step_begin < 0 || step_end < step_begin)
        return fail(OEM_EINVAL, "sampling window [%d, %d) every %d, %d slots", step_begin,
                    step_end, interval, n_slots);
    const int n[3] = {ctx->dims.nx, ctx->dims.ny, ctx->dims.nz};
    TimeMonDev d{};
    d.nc = n_comps;
    for (int t = 0; t < n_comps; ++t) {
        if (comps[t] < 0 || comps[t] > 5) return fail(OEM_EINVAL, "component %d", comps[t]);
        d.comps[t] = comps[t];
    }
    size_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (origin[a] < 0 || box[a] < 1 || origin[a] + box[a] > n[a])
            return fail(OEM_EINVAL, "box outside the grid on axis %d", a);
        d.o[a] = origin[a];
        d.b[a] = box[a];
        cells *= (size_t)box[a];
    }
    const size_t count = (size_t)n_comps * n_slots * cells;
    if (cells > ((size_t)1 << 30) || count > ((size_t)1 << 34))
        return fail(OEM_EINVAL, "time monitor too large (%zu values)", count);
    d.cells = (int)cells;
    d.beg = step_begin;
    d.end = step_end;
    d.iv = interval;
    d.slots = n_slots;
    float *buf = nullptr;
    TRY(ctx->alloc(&buf, count * ctx->esz, true));
    d.buf = buf;
    ctx->tmons.push_back(TimeMonitor{d, count});
    *mon_id = (int32_t)ctx->tmons.size() - 1;
    return OEM_OK;
}

extern "C" int oem_read_time_monitor(oem_ctx *ctx, int32_t mon_id, float *out, size_t count) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->tmons.size()) return fail(OEM_EINVAL, "time monitor %d", mon_id);
    const TimeMonitor &m = ctx->tmons[mon_id];
    if (!out || count != m.count * ctx->esz)
        return fail(OEM_EINVAL, "time monitor %d holds %zu floats, asked %zu", mon_id,
                    m.count * ctx->esz, count);
    HIP_OK(hipMemcpyAsync(out, m.d.buf, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_add_flux_time_monitor(
    oem_ctx *ctx, int32_t axis, int32_t n1, const int32_t *e1, const int32_t *l1,
    const int32_t *h1, const float *w1, const double *tr1, int32_t n2, const int32_t *e2,
    const int32_t *l2, const int32_t *h2, const float *w2, const double *tr2, int32_t n_planes,
    const double *plane_w, const int32_t *ke, const int32_t *kh, const float *wh_lo,
    const float *wh_hi, double sign, int32_t step_begin, int32_t step_end, int32_t interval,
    int32_t n_slots, int32_t *mon_id) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return fail(OEM_ESTATE, "context already prepared");
    if (ctx->cplx) return fail(OEM_ENOTSUP, "flux-time monitors on real fields only");
    if (!mon_id || !plane_w || !ke || !kh || !wh_lo || !wh_hi) return fail(OEM_EINVAL, "null argument");
    if (axis < 0 || axis > 2) return fail(OEM_EINVAL, "axis %d", axis);
    if (n1 < 1 || n2 < 1 || (int64_t)n1 * n2 > (1 << 28)) return fail(OEM_EINVAL, "%d x %d samples", n1, n2);
    if (n_planes < 1 || n_planes > 2) return fail(OEM_EINVAL, "%d planes", n_planes);
    if (interval < 1 || n_slots < 1 || step_begin < 0 || step_end < step_begin)
        return fail(OEM_EINVAL, "sampling window [%d, %d) every %d, %d slots", step_begin,
                    step_end, interval, n_slots);
    const int n[3] = {ctx->dims.nx, ctx->dims.ny, ctx->dims.nz};
    const int t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
    const int32_t *idx[2][3] = {{e1, l1, h1}, {e2, l2, h2}};
    const int cnt[2] = {n1, n2}, lim[2] = {n[t1], n[t2]};
    for (int t = 0; t < 2; ++t)
        for (int f = 0; f < 3; ++f) {
            if (!idx[t][f]) return fail(OEM_EINVAL, "null sample table");
            for (int q = 0; q < cnt[t]; ++q)
                if (idx[t][f][q] < 0 || idx[t][f][q] >= lim[t])
                    return fail(OEM_EINVAL, "transverse sample index out of range");
        }
    if (!w1 || !w2 || !tr1 || !tr2) return fail(OEM_EINVAL, "null weights");
    for (int p = 0; p < n_planes; ++p)
        if (ke[p] < 0 || ke[p] >= n[axis] || kh[p] < 0 || kh[p] + 1 >= n[axis])
            return fail(OEM_EINVAL, "normal plane %d out of range", p);
    FluxTimeDev d{};
    d.axis = axis; d.t1 = t1; d.t2 = t2; d.n1 = n1; d.n2 = n2; d.n_planes = n_planes;
    d.sign = sign; d.beg = step_begin; d.end = step_end; d.iv = interval; d.slots = n_slots;
    for (int p = 0; p < n_planes; ++p) {
        d.plane_w[p] = plane_w[p]; d.ke[p] = ke[p]; d.kh[p] = kh[p];
        d.wh_lo[p] = wh_lo[p]; d.wh_hi[p] = wh_hi[p];
    }
    int *di[2][3];
    for (int t = 0; t < 2; ++t)
        for (int f = 0; f < 3; ++f) TRY(ctx->upload(&di[t][f], (const int *)idx[t][f], (size_t)cnt[t]));
    d.e1 = di[0][0]; d.l1 = di[0][1]; d.h1 = di[0][2];
    d.e2 = di[1][0]; d.l2 = di[1][1]; d.h2 = di[1][2];
    float *dw1, *dw2;
    TRY(ctx->upload(&dw1, w1, (size_t)n1));
    TRY(ctx->upload(&dw2, w2, (size_t)n2));
    d.w1 = dw1; d.w2 = dw2;
    double *dt1, *dt2;
    TRY(ctx->upload(&dt1, tr1, (size_t)n1));
    TRY(ctx->upload(&dt2, tr2, (size_t)n2));
    d.tr1 = dt1; d.tr2 = dt2;
    const size_t np = (size_t)n1 * n2;
    d.n_blocks = (int)((np + OEM_FT_BLK - 1) / OEM_FT_BLK);
    float *ep;
    TRY(ctx->alloc(&ep, (size_t)2 * n_planes * np, true));
    d.ep = ep;
    double *part, *buf;
    TRY(ctx->alloc(&part, (size_t)2 * d.n_blocks, true));
    TRY(ctx->alloc(&buf, (size_t)n_slots, true));
    d.partial = part;
    d.buf = buf;
    ctx->ftmons.push_back(FluxTimeMonitor{d});
    *mon_id = (int32_t)ctx->ftmons.size() - 1;
    return OEM_OK;
}

extern "C" int oem_read_flux_time_monitor(oem_ctx *ctx, int32_t mon_id, double *out, size_t count) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->ftmons.size()) return fail(OEM_EINVAL, "flux-time monitor %d", mon_id);
    const FluxTimeMonitor &m = ctx->ftmons[mon_id];
    if (!out || count != (size_t)m.d.slots)
        return fail(OEM_EINVAL, "flux-time monitor %d holds %d values, asked %zu", mon_id, m.d.slots, count);
    HIP_OK(hipMemcpyAsync(out, m.d.buf, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

// The DFT monitors whose snapshots the fused step writes: up to OEM_FUSE_MONS, in order, as long
// as no work-group (footprints in x chunks x row tiles x k tiles) meets more than three of them
// (real fields) or more than one (complex fields: the complex march's extra stores cost more
// than the snapshot kernel they replace); the others keep k_dft_snap.  Each comes with the list of its
// points at cells that an E-side point list corrects after the step (those are snapped again
// after the corrections, so every snapshot holds the final fields).
static int plan_fused_snaps(oem_ctx *ctx) {
    const int nx = ctx->dims.nx, ny = ctx->dims.ny, nz = ctx->dims.nz;
    const int tile[3] = {ctx->fg.chunk, OEM_CR, 64};
    std::vector<std::array<int, 6>> taken;   // fused footprints: first, last tile per axis
    auto footprint = [&](const MonDev &m) {
        int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        if (m.kind == 1) {
            for (int a = 0; a < 3; ++a) { lo[a] = m.o[a]; hi[a] = m.o[a] + m.b[a]; }
        } else {
            lo[m.axis] = m.km;
            hi[m.axis] = m.km + 1;
        }
        std::array<int, 6> f{};
        for (int a = 0; a < 3; ++a) { f[2 * a] = lo[a] / tile[a]; f[2 * a + 1] = (hi[a] - 1) / tile[a]; }
        return f;
    };
    std::vector<int> ecells;
    for (auto &pt : ctx->points)
        if (!pt.magnetic) ecells.insert(ecells.end(), pt.cells.begin(), pt.cells.end());
    std::sort(ecells.begin(), ecells.end());
    ecells.erase(std::unique(ecells.begin(), ecells.end()), ecells.end());
    ctx->fsnap.n = 0;
    ctx->fsnap_multi = false;
    for (auto &mon : ctx->mons) {
        if (ctx->fsnap.n == OEM_FUSE_MONS) break;
        const MonDev &m = mon.d;
        const std::array<int, 6> fp = footprint(m);
        // the footprint shared by two boxes of tiles (first > last on an axis: none)
        auto meet = [](const std::array<int, 6> &a, const std::array<int, 6> &b) {
            std::array<int, 6> c{};
            for (int q = 0; q < 3; ++q) {
                c[2 * q] = std::max(a[2 * q], b[2 * q]);
                c[2 * q + 1] = std::min(a[2 * q + 1], b[2 * q + 1]);
            }
            return c;
        };
        auto empty = [](const std::array<int, 6> &c) {
            return c[0] > c[1] || c[2] > c[3] || c[4] > c[5];
        };
        // a second (complex) or fourth (real) in any work-group: the candidate and one or three
        // taken footprints share a tile
        bool shared = false, fourth = false;
        for (size_t a = 0; a < taken.size() && !fourth; ++a) {
            const std::array<int, 6> fa = meet(fp, taken[a]);
            if (empty(fa)) continue;
            shared = true;
            for (size_t b = a + 1; b < taken.size() && !fourth; ++b) {
                const std::array<int, 6> fb = meet(fa, taken[b]);
                if (empty(fb)) continue;
                for (size_t c = b + 1; c < taken.size() && !fourth; ++c)
                    fourth = !empty(meet(fb, taken[c]));
            }
        }
        if (fourth || (shared && ctx->cplx)) continue;
        ctx->fsnap_multi |= shared;
        taken.push_back(fp);
        std::vector<int> qs;
        for (const int id : ecells) {
            const int i = id / (ny * nz), r = id - i * ny * nz, j = r / nz, k = r - j * nz;
            if (m.kind == 1) {
                if (i < m.o[0] || i >= m.o[0] + m.b[0] || j < m.o[1] || j >= m.o[1] + m.b[1] ||
                    k < m.o[2] || k >= m.o[2] + m.b[2])
                    continue;
                qs.push_back(((i - m.o[0]) * m.b[1] + (j - m.o[1])) * m.b[2] + (k - m.o[2]));
            } else {
                if ((m.axis == 0 ? i : m.axis == 1 ? j : k) != m.km) continue;
                qs.push_back(m.axis == 0 ? j * nz + k : m.axis == 1 ? i * nz + k : i * ny + j);
            }
        }
        FuseMon &f = ctx->fsnap.m[ctx->fsnap.n++];
        f.kind = m.kind; f.axis = m.axis; f.km = m.km;
        for (int a = 0; a < 3; ++a) { f.o[a] = m.o[a]; f.b[a] = m.b[a]; }
        f.npts = m.npts; f.T = m.T; f.snap = m.snap;
        mon.fused = true;
        if (!qs.empty()) {
            TRY(ctx->upload(&mon.patch, qs.data(), qs.size()));
            mon.n_patch = (int)qs.size();
        }
    }
    return OEM_OK;
}

extern "C" int oem_prepare(oem_ctx *ctx) {
    CHECK_CTX(ctx);
    if (ctx->prepared) return OEM_OK;
    for (int a = 0; a < 3; ++a)
        if (!ctx->axis_set[a]) return fail(OEM_ESTATE, "axis %d tables not set", a);
    if (!ctx->mat_set) return fail(OEM_ESTATE, "materials not set");
    if (ctx->mat.mode == 0) {
        // every cell whose class reads the ADE history must lie inside the history box
        bool any = false;
        for (char f : ctx->class_flagged) any |= f != 0;
        if (any) {
            if (!ctx->has_disp) return fail(OEM_EINVAL, "classes flag dispersion but no ADE tables set");
            std::vector<uint16_t> cls(ctx->ncell);
            HIP_OK(hipMemcpyAsync(cls.data(), ctx->mat.cls, ctx->ncell * sizeof(uint16_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIP_OK(hipStreamSynchronize(ctx->stream));
            const int ny = ctx->dims.ny, nz = ctx->dims.nz;
            const int *hb = ctx->mat.hb;
            for (size_t id = 0; id < ctx->ncell; ++id) {
                if (!ctx->class_flagged[cls[id]]) continue;
                const int i = (int)(id / ((size_t)ny * nz));
                const int r = (int)(id - (size_t)i * ny * nz), j = r / nz, k = r % nz;
                if (i < hb[0] || i >= hb[1] || j < hb[2] || j >= hb[3] || k < hb[4] || k >= hb[5])
                    return fail(OEM_EINVAL, "cell %zu reads the ADE history outside its box", id);
            }
        }
    }
    GridDev &g = ctx->g;
    g.nx = ctx->dims.nx;
    g.ny = ctx->dims.ny;
    g.nz = ctx->dims.nz;
    g.plane = g.ny * g.nz;
    if ((size_t)g.plane * 4 >= ((size_t)1 << 32))
        return fail(OEM_EINVAL, "plane ny*nz=%d too large for 32-bit in-plane offsets", g.plane);
    g.flat = g.nz % OEM_TK != 0;
    g.tiles = g.flat ? (g.plane + OEM_TK * OEM_TJ - 1) / (OEM_TK * OEM_TJ)
                     : ((g.nz + OEM_TK - 1) / OEM_TK) * ((g.ny + OEM_TJ - 1) / OEM_TJ);
    int chunk;
    const char *env = std::getenv("OEM_CHUNK");
    if (ctx->chunk_req > 0) {
        chunk = ctx->chunk_req;
    } else if (env && std::atoi(env) > 0) {
        chunk = std::atoi(env);
    } else if (ctx->pool) {
        // a pool work-group marches a fixed number of planes: the batch of scenes supplies the
        // parallelism.  Pool throughput of k box soups of side n (Mcell-steps/s, chunks 4 / 6 /
        // 8 / 12 / 16 / 32 / 64): 48 x 64^3 26.9 / 27.0 / 26.8-27.0 / 26.3 / 25.9 / 26.3 / 23.6
        // G, 24 x 96^3 34.8 / 34.8 / 34.8-34.9 / 34.3 / 33.9 / 32.3 / 31.1 G, 12 x 128^3 37.0 /
        // 37.4 / 37.7 / 37.5 / 36.7 / 35.5 / 33.9 G, 6 x 160^3 42.3 / 43.2 / 43.1 / 42.7 / 42.5 /
        // 40.4 / 34.5 G, 3 x 180^3 42.3 / 42.5 / 42.7-42.8 / 41.2 / 40.0 / 36.6 / 32.0 G
        const char *pe = std::getenv("OEM_POOL_CHUNK");
        chunk = (pe && std::atoi(pe) > 0) ? std::atoi(pe) : 8;
    } else {
        // 64 x 8 tiles marching 32 planes at 512^3: measured best of tiles 4/8/16 rows and
        // chunks 32/64/128 planes
        const int want_blocks = 8192;
        const int chunks_needed = std::max(1, (want_blocks + g.tiles - 1) / g.tiles);
        chunk = std::max(4, (g.nx + chunks_needed - 1) / chunks_needed);
    }
    g.chunk = std::min(chunk, g.nx);
    const long long nb = (long long)g.tiles * ((g.nx + g.chunk - 1) / g.chunk);
    if (nb > 0x7fffffffLL) return fail(OEM_EINVAL, "launch too large");
    g.nblocks = (int)nb;
    // the fused H+E step for scenes of OEM_FUSE_MIN_CELLS and more (OEM_FUSE=1: any size, 0:
    // never), outside a serving pool.  Measured step times (box soup, two sweeps / fused):
    // 64^3 30 / 29-55 us, 128^3 69 / 65-83 us, 160^3 98 / 132 us, 200^3 241 / 240 us, 256^3
    // 393 / 355-361 us, 320^3 677 / 596 us, 512^3 2.54 / 1.99 ms: below ~8 M cells a fused
    // work-group marches too few tiles to fill the GPU
    // (oem_set_step_form overrides the rule and OEM_FUSE)
    const char *fe = std::getenv("OEM_FUSE");
    bool fuse = !ctx->pool && !(fe && fe[0] == '0');
    if (!(fe && fe[0] == '1')) fuse &= ctx->ncell >= OEM_FUSE_MIN_CELLS;
    if (ctx->form_req) fuse = ctx->form_req == 2;
    // it addresses a whole field set (E and H, six components) and a whole psi set (both sides)
    // with 32-bit offsets
    const bool fits = 6ull * ctx->ncell * 4ull * (uint64_t)ctx->esz < (1ull << 32) &&
                      2ull * ctx->psi.total * 4ull * (uint64_t)ctx->esz < (1ull << 32);
    if (ctx->form_req == 2 && !fits)
        return fail(OEM_ENOTSUP, "the fused step addresses a field set with 32-bit offsets");
    fuse &= fits;
    if (fuse) {
        const size_t nc = ctx->ncell;
        TRY(ctx->alloc(&ctx->E2, 6 * nc * ctx->esz, true));
        ctx->H2 = ctx->E2 + 3 * nc * ctx->esz;
        float *pp = nullptr;
        TRY(ctx->alloc(&pp, 2 * (size_t)ctx->psi.total * ctx->esz, true));
        ctx->psi2 = ctx->psi;
        ctx->psi2.h = pp;
        ctx->psi2.e = pp + (size_t)ctx->psi.total * ctx->esz;
        HIP_OK(hipStreamSynchronize(ctx->stream));
        // one work-group per (OEM_CR rows x 64 k) tile and x chunk.  Measured step times:
        // real fields at 512^3 (3 work-groups per CU), chunks 8 / 10 / 12 / 16 / 24 / 32: 2.00 /
        // 1.96 / 1.93 / 1.94 / 1.93 / 1.94 ms; the complex Bloch cell (2 per CU), chunks 12 / 16
        // / 24 / 32: 1.24 / 1.20 / 1.22 / 1.20 ms
        ctx->fg = g;
        ctx->fg.tiles = ((g.nz + 63) / 64) * ((g.ny + OEM_CR - 1) / OEM_CR);
        int fchunk = ctx->cplx ? 16 : 12;
        if (ctx->chunk_req > 0) fchunk = ctx->chunk_req;
        else if (env && std::atoi(env) > 0) fchunk = std::atoi(env);
        ctx->fg.chunk = std::min(fchunk, g.nx);
        const long long fb = (long long)ctx->fg.tiles * ((g.nx + ctx->fg.chunk - 1) / ctx->fg.chunk);
        if (fb > 0x7fffffffLL) return fail(OEM_EINVAL, "launch too large");
        ctx->fg.nblocks = (int)fb;
        ctx->fused = true;
        const char *fsn = std::getenv("OEM_FUSE_SNAP");
        if (!(fsn && fsn[0] == '0')) TRY(plan_fused_snaps(ctx));
    }
    ctx->prepared = true;
    return OEM_OK;
}

namespace {

struct Prof {
    hipEvent_t ev[64][2];
    int kid[64];
    int n = 0;
};

inline int blocks_for(int n, int bs) { return (n + bs - 1) / bs; }

template <typename F>
int launch_flush_t(oem_ctx *c, int cnt) {
    for (auto &m : c->mons) {
        const int ngrp = (m.d.nf + OEM_FLUSH_FG - 1) / OEM_FLUSH_FG;
        const dim3 grid(blocks_for(m.d.npts, 256), m.d.ncomp * ngrp);
        hipLaunchKernelGGL((k_dft_flush<F>), grid, dim3(256), 0, c->stream, m.d, cnt, c->d_step);
    }
    HIP_OK(hipGetLastError());
    return OEM_OK;
}

int launch_flush(oem_ctx *c, int cnt) {
    return c->cplx ? launch_flush_t<cf2>(c, cnt) : launch_flush_t<float>(c, cnt);
}

// Whether a step can end with k_step_tail, and its lists: the E-side point lists in step order
// and the monitor points they correct that the fused step snapped before the correction.  Not
// when anything else follows the E-side points and reads the step counter (an adjoint record,
// time monitors, monitors the snapshot kernel takes) or a list exceeds the block.
// OEM_TAIL=0 keeps the separate launches.
static bool tail_plan(const oem_ctx *c, TailDev *t) {
    const char *te = std::getenv("OEM_TAIL");
    if ((te && te[0] == '0') || c->rec_hook.on) return false;
    if (!c->mons_off && (!c->tmons.empty() || !c->ftmons.empty())) return false;
    t->n_pts = t->n_snap = 0;
    for (int slot = 3; slot < 6; ++slot)
        for (const auto &pt : c->points) {
            if (points_slot(pt) != slot) continue;
            if (t->n_pts == OEM_TAIL_LISTS || pt.d.n_groups > OEM_TAIL_THREADS) return false;
            t->pts[t->n_pts++] = pt.d;
        }
    if (!c->mons_off)
        for (const auto &m : c->mons) {
            if (!(c->fused && m.fused)) return false;
            if (!m.n_patch) continue;
            if (t->n_snap == OEM_TAIL_LISTS || m.n_patch > OEM_TAIL_THREADS) return false;
            t->mon[t->n_snap] = m.d;
            t->qs[t->n_snap] = m.patch;
            t->nq[t->n_snap++] = m.n_patch;
        }
    return true;
}

// One time step.  `flush_after`: number of snapshots to flush after the step counter advanced
// (0 = none).  `last`: the final step of a run call ends the ADE with a plain post, otherwise
// the post is fused with the next step's pre.  With `pr`, every launch is bracketed by events.
template <typename F>
int launch_step_t(oem_ctx *c, int flush_after, Prof *pr, bool last) {
    // the fields this step produces: the other set on the fused path
    const bool odd = c->fused && (c->step & 1);
    F *E = reinterpret_cast<F *>(c->fused && !odd ? c->E2 : c->E);
    F *H = reinterpret_cast<F *>(c->fused && !odd ? c->H2 : c->H);
    const size_t nc = c->ncell;
    auto mark = [&](int kid, bool begin) -> int {
        if (!pr) return OEM_OK;
        if (begin) {
            if (pr->n >= 64) return fail(OEM_EINVAL, "too many launches to profile");
            pr->kid[pr->n] = kid;
            HIP_OK(hipEventRecord(pr->ev[pr->n][0], c->stream));
        } else {
            HIP_OK(hipEventRecord(pr->ev[pr->n][1], c->stream));
            pr->n++;
        }
        return OEM_OK;
    };
    const dim3 yb(OEM_TK, OEM_TJ), yg(c->g.nblocks);
    const dim3 db(256), dg(c->has_disp ? blocks_for(c->disp.n_entry, 256) : 1);
    if (c->fused) {
        // ADE pre on the old E, then H and E in one march from the old set into the new one
        F *Eo = reinterpret_cast<F *>(odd ? c->E2 : c->E), *Ho = reinterpret_cast<F *>(odd ? c->H2 : c->H);
        const PsiDev &po = odd ? c->psi2 : c->psi, &pn = odd ? c->psi : c->psi2;
        if (c->has_disp && !c->pre_ready) {
            TRY(mark(OEM_K_DISP_PRE, true));
            hipLaunchKernelGGL((k_disp<F>), dg, db, 0, c->stream, c->disp, (const F *)Eo, nc,
                               c->dims.ny, c->dims.nz, 0);
            TRY(mark(OEM_K_DISP_PRE, false));
        }
        // H-side point corrections (magnetic dipoles, TF/SF faces, mode sources) are additive
        // terms independent of the fields: added to the old H before the march, every H update
        // and halo recompute sees them (the sweeps add them after update_h instead, which
        // rounds (H - ch curl E) + c rather than (H + c) - ch curl E)
        for (int slot = 0; slot < 3; ++slot)
            for (auto &pt : c->points) {
                if (points_slot(pt) != slot) continue;
                TRY(mark(points_kid(pt), true));
                hipLaunchKernelGGL((k_points<F>), dim3(blocks_for(pt.d.n_groups, 256)), dim3(256),
                                   0, c->stream, Ho, pt.d, c->mat, c->dims.ch, nc, c->d_step);
                TRY(mark(points_kid(pt), false));
            }
        TRY(mark(OEM_K_UPDATE_HE, true));
        const FuseSnap fsn = c->mons_off ? FuseSnap{} : c->fsnap;
        const dim3 hg(c->fg.nblocks), hb(64, OEM_CR + 1);
        const int mk = c->mat.mode != 0 ? 2 : c->mat.n_classes <= OEM_LDS_CLASSES ? 0 : 1;
#define OEM_HE_LAUNCH(MKV, MULTIV)                                                        \
        hipLaunchKernelGGL((k_update_he<F, MKV, MULTIV>), hg, hb, 0, c->stream,            \
                           (const F *)Eo, (const F *)Ho, E, H, po, pn, c->fg, c->mat,       \
                           c->dims.ch, c->dims.cw, c->pw, c->d_step, fsn)
        if (!c->mons_off && c->fsnap_multi) {
            if (mk == 0) OEM_HE_LAUNCH(0, true);
            else if (mk == 1) OEM_HE_LAUNCH(1, true);
            else OEM_HE_LAUNCH(2, true);
        } else {
            if (mk == 0) OEM_HE_LAUNCH(0, false);
            else if (mk == 1) OEM_HE_LAUNCH(1, false);
            else OEM_HE_LAUNCH(2, false);
        }
#undef OEM_HE_LAUNCH
        TRY(mark(OEM_K_UPDATE_HE, false));
    } else {
    TRY(mark(OEM_K_UPDATE_H, true));
    hipLaunchKernelGGL((k_update_h<F>), yg, yb, 0, c->stream, H, (const F *)E, c->psi, c->g,
                       c->dims.ch, c->pw, c->d_step);
    TRY(mark(OEM_K_UPDATE_H, false));
    for (int slot = 0; slot < 3; ++slot)
        for (auto &pt : c->points) {
            if (points_slot(pt) != slot) continue;
            TRY(mark(points_kid(pt), true));
            hipLaunchKernelGGL((k_points<F>), dim3(blocks_for(pt.d.n_groups, 256)), dim3(256), 0,
                               c->stream, H, pt.d, c->mat, c->dims.ch, nc, c->d_step);
            TRY(mark(points_kid(pt), false));
        }
    if (c->has_disp && !c->pre_ready) {
        TRY(mark(OEM_K_DISP_PRE, true));
        hipLaunchKernelGGL((k_disp<F>), dg, db, 0, c->stream, c->disp, (const F *)E, nc,
                           c->dims.ny, c->dims.nz, 0);
        TRY(mark(OEM_K_DISP_PRE, false));
    }
    TRY(mark(OEM_K_UPDATE_E, true));
    hipLaunchKernelGGL((k_update_e<F>), yg, yb, 0, c->stream, E, (const F *)H, c->psi, c->g,
                       c->mat, c->dims.cw, c->pw, c->d_step);
    TRY(mark(OEM_K_UPDATE_E, false));
    }
    // ADE post on the new E (after every E-side correction); the post of a run's last step is
    // plain, any other is fused with the next step's pre
    auto disp_post = [&]() -> int {
        if (!c->has_disp) return OEM_OK;
        TRY(mark(OEM_K_DISP_POST, true));
        hipLaunchKernelGGL((k_disp<F>), dg, db, 0, c->stream, c->disp, (const F *)E, nc,
                           c->dims.ny, c->dims.nz, last ? 1 : 2);
        TRY(mark(OEM_K_DISP_POST, false));
        c->pre_ready = !last;
        return OEM_OK;
    };
    TailDev tail{};
    if (tail_plan(c, &tail)) {
        // one launch: the E-side points, the monitor points they corrected snapped again, the
        // step counter; then the ADE post, which reads no step counter
        TRY(mark(OEM_K_STEP_TAIL, true));
        hipLaunchKernelGGL((k_step_tail<F>), dim3(1), dim3(OEM_TAIL_THREADS), 0, c->stream, E,
                           (const F *)H, tail, c->mat, c->dims.ch, nc, c->dims.ny, c->dims.nz,
                           c->d_step);
        TRY(mark(OEM_K_STEP_TAIL, false));
        TRY(disp_post());
    } else {
        // E-side corrections: TF/SF faces, mode sources, then electric dipoles (reference order,
        // `openem/solver.py:945-968`)
        for (int slot = 3; slot < 6; ++slot)
            for (auto &pt : c->points) {
                if (points_slot(pt) != slot) continue;
                TRY(mark(points_kid(pt), true));
                hipLaunchKernelGGL((k_points<F>), dim3(blocks_for(pt.d.n_groups, 256)), dim3(256), 0,
                                   c->stream, E, pt.d, c->mat, c->dims.ch, nc, c->d_step);
                TRY(mark(points_kid(pt), false));
            }
        TRY(disp_post());
        if (c->rec_hook.on)
            hipLaunchKernelGGL((k_rec_gather<F>), dim3(blocks_for(c->rec_hook.n, 256)), dim3(256), 0,
                               c->stream, (const F *)E, c->rec_hook.comp, c->rec_hook.cell,
                               c->rec_hook.n, nc, reinterpret_cast<F *>(c->rec_hook.rec), c->d_step,
                               c->rec_hook.base);
        for (auto &tm : c->tmons) {
            if (c->mons_off) break;
            for (int shot = 0; shot < 2; ++shot) {
                TRY(mark(OEM_K_DFT_SNAP, true));
                hipLaunchKernelGGL((k_time_sample<F>), dim3(blocks_for(tm.d.cells, 256)), dim3(256), 0,
                                   c->stream, tm.d, (const F *)E, (const F *)H, nc, c->dims.ny,
                                   c->dims.nz, c->d_step, shot);
                TRY(mark(OEM_K_DFT_SNAP, false));
            }
        }
        if constexpr (sizeof(F) == 4) {
            for (auto &fm : c->ftmons) {
                if (c->mons_off) break;
                TRY(mark(OEM_K_DFT_SNAP, true));
                hipLaunchKernelGGL((k_flux_time_partial<float>), dim3(fm.d.n_blocks), dim3(OEM_FT_BLK), 0,
                                   c->stream, fm.d, (const float *)E, (const float *)H, nc,
                                   c->dims.ny, c->dims.nz, c->d_step);
                hipLaunchKernelGGL((k_flux_time_sum<float>), dim3(1), dim3(OEM_FT_BLK), 0, c->stream, fm.d,
                                   c->d_step);
                TRY(mark(OEM_K_DFT_SNAP, false));
            }
        }
        for (auto &m : c->mons) {
            if (c->mons_off) break;
            if (c->fused && m.fused && !m.n_patch) continue;
            TRY(mark(OEM_K_DFT_SNAP, true));
            if (c->fused && m.fused)
                hipLaunchKernelGGL((k_dft_snap_list<F>), dim3(blocks_for(m.n_patch, 256)), dim3(256), 0,
                                   c->stream, m.d, (const F *)E, (const F *)H, nc, c->dims.ny,
                                   c->dims.nz, c->d_step, (const int *)m.patch, m.n_patch);
            else
                hipLaunchKernelGGL((k_dft_snap<F>), dim3(blocks_for(m.d.npts, 256)), dim3(256), 0,
                                   c->stream, m.d, (const F *)E, (const F *)H, nc, c->dims.ny,
                                   c->dims.nz, c->d_step);
            TRY(mark(OEM_K_DFT_SNAP, false));
        }
        TRY(mark(OEM_K_STEP, true));
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, c->stream, c->d_step);
        TRY(mark(OEM_K_STEP, false));
    }
    if (flush_after > 0 && !c->mons.empty() && !c->mons_off) {
        TRY(mark(OEM_K_DFT_FLUSH, true));
        TRY(launch_flush(c, flush_after));
        TRY(mark(OEM_K_DFT_FLUSH, false));
    }
    HIP_OK(hipGetLastError());
    return OEM_OK;
}

int launch_step(oem_ctx *c, int flush_after, Prof *pr, bool last) {
    return c->cplx ? launch_step_t<cf2>(c, flush_after, pr, last)
                   : launch_step_t<float>(c, flush_after, pr, last);
}

}  // namespace

int oemi::launch_one_step(oem_ctx *ctx, bool last) {
    TRY(launch_step(ctx, 0, nullptr, last));
    ctx->step++;
    return OEM_OK;
}

static int check_run(oem_ctx *ctx, int64_t n) {
    if (!ctx->prepared) return fail(OEM_ESTATE, "oem_prepare not called");
    if (ctx->attached) return fail(OEM_ESTATE, "context is stepped by its pool");
    if (n < 0) return fail(OEM_EINVAL, "negative step count");
    if (ctx->step + n > ctx->dims.n_steps)
        return fail(OEM_EINVAL, "step %lld + %lld exceeds the %d-step tables",
                    (long long)ctx->step, (long long)n, ctx->dims.n_steps);
    return OEM_OK;
}

static int run_steps(oem_ctx *ctx, int64_t n) {
    const int T = ctx->dims.dft_batch;
    const bool mons = !ctx->mons.empty();
    const char *ng = std::getenv("OEM_NO_GRAPH");
    const bool use_graph = !(ng && ng[0] == '1');
    int64_t left = n;
    while (left > 0) {
        const bool disp_ok = !ctx->has_disp || ctx->pre_ready;
        if (use_graph && ctx->pending == 0 && ctx->step % T == 0 && left > T && disp_ok &&
            !(ctx->fused && (T & 1))) {
            if (!ctx->graph) {
                // no other thread synchronises the device while this stream captures
                std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
                hipGraph_t gr = nullptr;
                HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                // each captured step sees its own step index: on the fused path its parity picks
                // which field set is old and which is new
                int rc = OEM_OK;
                const int64_t s0 = ctx->step;
                for (int s = 0; s < T && !rc; ++s) {
                    ctx->step = s0 + s;
                    rc = launch_step(ctx, s == T - 1 ? T : 0, nullptr, false);
                }
                ctx->step = s0;
                hipError_t ec = hipStreamEndCapture(ctx->stream, &gr);
                if (rc) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return rc;
                }
                if (ec != hipSuccess) return fail(OEM_EHIP, "graph capture: %s", hipGetErrorString(ec));
                hipError_t ei = hipGraphInstantiate(&ctx->graph, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (ei != hipSuccess) {
                    ctx->graph = nullptr;
                    return fail(OEM_EHIP, "graph instantiate: %s", hipGetErrorString(ei));
                }
            }
            HIP_OK(hipGraphLaunch(ctx->graph, ctx->stream));
            ctx->step += T;
            left -= T;
        } else {
            int fl = 0;
            if (mons) {
                ctx->pending++;
                if (ctx->pending == T) fl = T;
            }
            TRY(launch_step(ctx, fl, nullptr, left == 1));
            if (fl) ctx->pending = 0;
            ctx->step++;
            left--;
        }
    }
    if (mons && ctx->pending > 0) {
        TRY(launch_flush(ctx, ctx->pending));
        ctx->pending = 0;
    }
    return OEM_OK;
}

extern "C" int oem_run(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(check_run(ctx, n_steps));
    TRY(run_steps(ctx, n_steps));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_timed_run(oem_ctx *ctx, int64_t n_steps, double *ms) {
    CHECK_CTX(ctx);
    if (!ms) return fail(OEM_EINVAL, "null output");
    TRY(check_run(ctx, n_steps));
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    HIP_OK(hipEventRecord(a, ctx->stream));
    int rc = run_steps(ctx, n_steps);
    if (!rc) {
        HIP_OK(hipEventRecord(b, ctx->stream));
        HIP_OK(hipEventSynchronize(b));
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, a, b));
        *ms = t;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

extern "C" int oem_profile_run(oem_ctx *ctx, int64_t n_steps, double *ms, int64_t *calls) {
    CHECK_CTX(ctx);
    if (!ms || !calls) return fail(OEM_EINVAL, "null output");
    TRY(check_run(ctx, n_steps));
    for (int k = 0; k < OEM_K_COUNT; ++k) { ms[k] = 0.0; calls[k] = 0; }
    Prof *pr = new Prof();
    int rc = OEM_OK;
    for (int q = 0; q < 64 && !rc; ++q)
        for (int e = 0; e < 2; ++e)
            if (hipEventCreate(&pr->ev[q][e]) != hipSuccess) rc = fail(OEM_EHIP, "event create");
    const int T = ctx->dims.dft_batch;
    const bool mons = !ctx->mons.empty();
    for (int64_t s = 0; s < n_steps && !rc; ++s) {
        pr->n = 0;
        int fl = 0;
        if (mons) {
            ctx->pending++;
            if (ctx->pending == T || s == n_steps - 1) fl = ctx->pending;
        }
        rc = launch_step(ctx, fl, pr, s == n_steps - 1);
        if (rc) break;
        if (fl) ctx->pending = 0;
        ctx->step++;
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) { rc = fail(OEM_EHIP, "sync"); break; }
        for (int q = 0; q < pr->n; ++q) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, pr->ev[q][0], pr->ev[q][1]);
            ms[pr->kid[q]] += t;
            calls[pr->kid[q]] += 1;
        }
    }
    for (int q = 0; q < 64; ++q)
        for (int e = 0; e < 2; ++e)
            if (pr->ev[q][e]) (void)hipEventDestroy(pr->ev[q][e]);
    delete pr;
    return rc;
}

extern "C" int oem_step_index(oem_ctx *ctx, int64_t *n) {
    if (!ctx || !n) return fail(OEM_EINVAL, "null argument");
    *n = ctx->step;
    return OEM_OK;
}

extern "C" int oem_read_monitor(oem_ctx *ctx, int32_t mon_id, double *re, double *im,
                                size_t count) {
    CHECK_CTX(ctx);
    if (mon_id < 0 || mon_id >= (int)ctx->mons.size()) return fail(OEM_EINVAL, "monitor %d", mon_id);
    const Monitor &m = ctx->mons[mon_id];
    if (count != m.count || !re || !im)
        return fail(OEM_EINVAL, "monitor %d holds %zu values, asked %zu", mon_id, m.count, count);
    if (ctx->pool) TRY(oemi::pool_wait_done(ctx));
    HIP_OK(hipMemcpyAsync(re, m.d.re, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(im, m.d.im, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

static int field_ptr(oem_ctx *ctx, int32_t comp, size_t count, float **p) {
    if (comp < 0 || comp > 5) return fail(OEM_EINVAL, "component %d", comp);
    if (count != ctx->ncell * ctx->esz)
        return fail(OEM_EINVAL, "field holds %zu floats, got %zu", ctx->ncell * ctx->esz, count);
    *p = (comp < 3 ? ctx->cur_E() : ctx->cur_H()) + (size_t)(comp % 3) * ctx->ncell * ctx->esz;
    return OEM_OK;
}

extern "C" int oem_read_field(oem_ctx *ctx, int32_t comp, float *out, size_t count) {
    CHECK_CTX(ctx);
    float *p;
    TRY(field_ptr(ctx, comp, count, &p));
    if (!out) return fail(OEM_EINVAL, "null output");
    if (ctx->pool) TRY(oemi::pool_wait_done(ctx));
    HIP_OK(hipMemcpyAsync(out, p, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_write_field(oem_ctx *ctx, int32_t comp, const float *in, size_t count) {
    CHECK_CTX(ctx);
    float *p;
    TRY(field_ptr(ctx, comp, count, &p));
    if (!in) return fail(OEM_EINVAL, "null input");
    if (ctx->attached) return fail(OEM_ESTATE, "context is stepped by its pool");
    HIP_OK(hipMemcpyAsync(p, in, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}
