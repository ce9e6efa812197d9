// Device kernels of the MI355X (gfx950, CDNA4) FDTD stepper.
//
// The Yee update is HBM-bound (about 1 flop per byte), so the design goal is: every field
// byte crosses HBM once per sweep.  Each work-group owns a tile of 256 consecutive positions
// p = j*nz + k of the (ny, nz) plane and marches along x (the slowest axis) over a chunk of
// i-planes.  The x neighbour plane that the curl needs is carried in registers from one
// iteration to the next; the y/z neighbours are in the same or an adjacent plane tile and are
// served by L1/L2.  The linear block id is remapped so that consecutive tiles land on the same
// XCD (MI355X dispatches work-groups round-robin over 8 XCDs, each with its own L2), which keeps
// the halo rows of neighbouring tiles in one L2.
//
// CPML psi is stored only inside the PML slabs of each axis.  Outside a slab the reference's
// identity coefficients (a=0, b=1, 1/kappa=1, `openem/kernels/yee.cu:60-74`) make the stretched
// derivative equal to the plain one, which is what the kernels compute there.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OEM_TILE 256
#define OEM_NXCD 8
#define OEM_MAX_PW 2

struct AxisDev {
    const float *idl_p, *idl_d;
    const int *nxt, *prv;
    const float *mnx, *mpv, *pec;
    const float *a_e, *b_e, *k_e;
    const float *a_h, *b_h, *k_h;
    int n, lo, hi;
};

struct GridDev {
    AxisDev ax[3];
    int nx, ny, nz;
    int plane;          // ny*nz
    int tiles;          // ceil(plane / OEM_TILE)
    int chunk;          // i-planes per block
    int nblocks;        // tiles * ceil(nx / chunk)
};

// psi slabs; index [2*axis + c] where c selects the field component pair member:
//   H side: x -> (Hy, Hz), y -> (Hx, Hz), z -> (Hx, Hy); E side likewise with E.
struct PsiDev {
    float *h[6];
    float *e[6];
};

struct PlaneSrcDev {
    int n;              // active plane sources (0..OEM_MAX_PW)
    int pol[OEM_MAX_PW], kh[OEM_MAX_PW], ks[OEM_MAX_PW];
    float coef_h[OEM_MAX_PW], coef_e[OEM_MAX_PW];
    const double *ex_inc[OEM_MAX_PW], *hy_inc[OEM_MAX_PW];
};

struct MatDev {
    int mode;           // 0: class LUT, 1: dense ca/cb
    const uint16_t *cls;
    const float *lut;   // n_classes * 8
    const float *ca, *cb;   // dense: 3*ncell each
    float *hist;        // (3, bx, by, bz) inside hist box, or nullptr
    int hb[6];          // x0 x1 y0 y1 z0 z1
};

// Round-robin XCD dispatch -> contiguous logical ranges per XCD (a bijection on [0, nb)).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb / OEM_NXCD, rem = nb % OEM_NXCD;
    const int x = b % OEM_NXCD, s = b / OEM_NXCD;
    return x < rem ? x * (per + 1) + s : rem * (per + 1) + (x - rem) * per + s;
}

__device__ __forceinline__ int slab_of(int i, int n, int lo, int hi) {
    return i < lo ? i : (i >= n - hi ? i - (n - hi) + lo : -1);
}

// CPML recursion on one psi element: psi <- b*psi + a*d; returns invk*d + psi.
__device__ __forceinline__ float stretch(float *psi, size_t id, float d, float a, float b,
                                         float invk) {
    const float p = b * psi[id] + a * d;
    psi[id] = p;
    return invk * d + p;
}

__device__ __forceinline__ int step_now(const int *stp) { return *stp; }

// ------------------------------------------------------------------------------------------
// H half step (reference update_h, `openem/kernels/yee.cu:76-137`) fused with the H side of
// the plane-wave TF/SF correction (`openem/kernels/source_dft.cu:40-55`).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(OEM_TILE)
k_update_h(float *__restrict__ Hx, float *__restrict__ Hy, float *__restrict__ Hz,
           const float *__restrict__ Ex, const float *__restrict__ Ey,
           const float *__restrict__ Ez, PsiDev ps, GridDev g, float ch, PlaneSrcDev pw,
           const int *__restrict__ stp)
{
    const int lb = xcd_remap(blockIdx.x, g.nblocks);
    const int tile = lb % g.tiles, chunk = lb / g.tiles;
    const int p = tile * OEM_TILE + threadIdx.x;
    if (p >= g.plane) return;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const int j = p / nz, k = p - j * nz;
    const int i0 = chunk * g.chunk;
    const int i1 = min(i0 + g.chunk, nx);
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];

    const int pj = Y.nxt[j] * nz + k;      // (j+1, k) inside the plane
    const int pk = j * nz + Z.nxt[k];      // (j, k+1)
    const float mj = Y.mnx[j], mk = Z.mnx[k];
    const float idy = Y.idl_p[j], idz = Z.idl_p[k];
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = Y.a_h[j]; by = Y.b_h[j]; ky = Y.k_h[j]; }
    if (sz >= 0) { az = Z.a_h[k]; bz = Z.b_h[k]; kz = Z.k_h[k]; }

    // plane-wave H correction of this column (k == kh), amplitude of this step
    float inj_hx = 0.f, inj_hy = 0.f;
    if (pw.n) {
        const int n = step_now(stp);
        for (int s = 0; s < pw.n; ++s) {
            if (k == pw.kh[s]) {
                const float v = pw.coef_h[s] * (float)pw.ex_inc[s][n];
                if (pw.pol[s] == 0) inj_hy += v; else inj_hx += v;
            }
        }
    }

    const size_t plane = (size_t)g.plane;
    float ey = Ey[(size_t)i0 * plane + p];
    float ez = Ez[(size_t)i0 * plane + p];
    for (int i = i0; i < i1; ++i) {
        const size_t base = (size_t)i * plane;
        const size_t id = base + p;
        const size_t idn = (size_t)X.nxt[i] * plane + p;
        const float mi = X.mnx[i], idx = X.idl_p[i];
        const float ex = Ex[id];
        const float ex_j = Ex[base + pj], ex_k = Ex[base + pk];
        const float ey_k = Ey[base + pk];
        const float ez_j = Ez[base + pj];
        const float ey_n = Ey[idn], ez_n = Ez[idn];

        const float dzy = (ez_j * mj - ez) * idy;
        const float dyz = (ey_k * mk - ey) * idz;
        const float dxz = (ex_k * mk - ex) * idz;
        const float dzx = (ez_n * mi - ez) * idx;
        const float dyx = (ey_n * mi - ey) * idx;
        const float dxy = (ex_j * mj - ex) * idy;

        float s_hx_y = dzy, s_hz_y = dxy, s_hx_z = dyz, s_hy_z = dxz, s_hy_x = dzx, s_hz_x = dyx;
        const int sx = slab_of(i, nx, X.lo, X.hi);
        if (sx >= 0) {
            const size_t q = ((size_t)sx * ny + j) * nz + k;
            const float ax_ = X.a_h[i], bx_ = X.b_h[i], kx_ = X.k_h[i];
            s_hy_x = stretch(ps.h[0], q, dzx, ax_, bx_, kx_);
            s_hz_x = stretch(ps.h[1], q, dyx, ax_, bx_, kx_);
        }
        if (sy >= 0) {
            const size_t q = ((size_t)i * wy + sy) * nz + k;
            s_hx_y = stretch(ps.h[2], q, dzy, ay, by, ky);
            s_hz_y = stretch(ps.h[3], q, dxy, ay, by, ky);
        }
        if (sz >= 0) {
            const size_t q = ((size_t)i * ny + j) * wz + sz;
            s_hx_z = stretch(ps.h[4], q, dyz, az, bz, kz);
            s_hy_z = stretch(ps.h[5], q, dxz, az, bz, kz);
        }
        Hx[id] = Hx[id] - ch * (s_hx_y - s_hx_z) + inj_hx;
        Hy[id] = Hy[id] - ch * (s_hy_z - s_hy_x) + inj_hy;
        Hz[id] = Hz[id] - ch * (s_hz_x - s_hz_y);
        ey = ey_n;
        ez = ez_n;
    }
}

// ------------------------------------------------------------------------------------------
// E whole step (reference update_e / update_e_lut, `openem/kernels/yee.cu:154-337`) fused with
// the E side of the plane-wave TF/SF correction (`openem/kernels/source_dft.cu:58-73`).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(OEM_TILE)
k_update_e(float *__restrict__ Ex, float *__restrict__ Ey, float *__restrict__ Ez,
           const float *__restrict__ Hx, const float *__restrict__ Hy,
           const float *__restrict__ Hz, PsiDev ps, GridDev g, MatDev mat, float cw,
           PlaneSrcDev pw, const int *__restrict__ stp)
{
    const int lb = xcd_remap(blockIdx.x, g.nblocks);
    const int tile = lb % g.tiles, chunk = lb / g.tiles;
    const int p = tile * OEM_TILE + threadIdx.x;
    if (p >= g.plane) return;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const int j = p / nz, k = p - j * nz;
    const int i0 = chunk * g.chunk;
    const int i1 = min(i0 + g.chunk, nx);
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];

    const int pj = Y.prv[j] * nz + k;      // (j-1, k)
    const int pk = j * nz + Z.prv[k];      // (j, k-1)
    const float mj = Y.mpv[j], mk = Z.mpv[k];
    const float idy = Y.idl_d[j], idz = Z.idl_d[k];
    const float pecy = Y.pec[j], pecz = Z.pec[k];
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = Y.a_e[j]; by = Y.b_e[j]; ky = Y.k_e[j]; }
    if (sz >= 0) { az = Z.a_e[k]; bz = Z.b_e[k]; kz = Z.k_e[k]; }

    float inj_ex = 0.f, inj_ey = 0.f;       // coef_e * Hy_inc of this column (k == ks)
    bool inj_x = false, inj_y = false;
    if (pw.n) {
        const int n = step_now(stp);
        for (int s = 0; s < pw.n; ++s) {
            if (k == pw.ks[s]) {
                const float v = pw.coef_e[s] * (float)pw.hy_inc[s][n];
                if (pw.pol[s] == 0) { inj_ex += v; inj_x = true; }
                else { inj_ey += v; inj_y = true; }
            }
        }
    }
    const bool hrow = mat.hist && j >= mat.hb[2] && j < mat.hb[3] && k >= mat.hb[4] &&
                      k < mat.hb[5];
    const int hbx = mat.hb[1] - mat.hb[0], hby = mat.hb[3] - mat.hb[2],
              hbz = mat.hb[5] - mat.hb[4];
    const size_t hcomp = (size_t)hbx * hby * hbz;

    const size_t plane = (size_t)g.plane;
    const size_t ncell = plane * nx;
    const size_t idp0 = (size_t)X.prv[i0] * plane + p;
    float hy_m = Hy[idp0], hz_m = Hz[idp0];
    for (int i = i0; i < i1; ++i) {
        const size_t base = (size_t)i * plane;
        const size_t id = base + p;
        const float mi = X.mpv[i], idx = X.idl_d[i], pecx = X.pec[i];
        const float hx = Hx[id], hy = Hy[id], hz = Hz[id];
        const float hx_j = Hx[base + pj], hx_k = Hx[base + pk];
        const float hy_k = Hy[base + pk];
        const float hz_j = Hz[base + pj];

        const float dzy = (hz - hz_j * mj) * idy;
        const float dyz = (hy - hy_k * mk) * idz;
        const float dxz = (hx - hx_k * mk) * idz;
        const float dzx = (hz - hz_m * mi) * idx;
        const float dyx = (hy - hy_m * mi) * idx;
        const float dxy = (hx - hx_j * mj) * idy;

        float s_ex_y = dzy, s_ez_y = dxy, s_ex_z = dyz, s_ey_z = dxz, s_ey_x = dzx, s_ez_x = dyx;
        const int sx = slab_of(i, nx, X.lo, X.hi);
        if (sx >= 0) {
            const size_t q = ((size_t)sx * ny + j) * nz + k;
            const float ax_ = X.a_e[i], bx_ = X.b_e[i], kx_ = X.k_e[i];
            s_ey_x = stretch(ps.e[0], q, dzx, ax_, bx_, kx_);
            s_ez_x = stretch(ps.e[1], q, dyx, ax_, bx_, kx_);
        }
        if (sy >= 0) {
            const size_t q = ((size_t)i * wy + sy) * nz + k;
            s_ex_y = stretch(ps.e[2], q, dzy, ay, by, ky);
            s_ez_y = stretch(ps.e[3], q, dxy, ay, by, ky);
        }
        if (sz >= 0) {
            const size_t q = ((size_t)i * ny + j) * wz + sz;
            s_ex_z = stretch(ps.e[4], q, dyz, az, bz, kz);
            s_ey_z = stretch(ps.e[5], q, dxz, az, bz, kz);
        }

        float cax, cbx, cay, cby, caz, cbz;
        uint32_t fl;
        if (mat.mode == 0) {
            const float4 *row = reinterpret_cast<const float4 *>(mat.lut + 8 * (size_t)mat.cls[id]);
            const float4 r0 = row[0], r1 = row[1];
            cax = r0.x; cbx = r0.y; cay = r0.z; cby = r0.w; caz = r1.x; cbz = r1.y;
            fl = __float_as_uint(r1.z);
        } else {
            cax = mat.ca[id]; cay = mat.ca[ncell + id]; caz = mat.ca[2 * ncell + id];
            cbx = mat.cb[id]; cby = mat.cb[ncell + id]; cbz = mat.cb[2 * ncell + id];
            fl = (hrow && i >= mat.hb[0] && i < mat.hb[1]) ? 7u : 0u;
        }
        float phx = 0.f, phy = 0.f, phz = 0.f;
        if (fl) {
            const size_t hq = ((size_t)(i - mat.hb[0]) * hby + (j - mat.hb[2])) * hbz +
                              (k - mat.hb[4]);
            if (fl & 1u) phx = mat.hist[hq];
            if (fl & 2u) phy = mat.hist[hcomp + hq];
            if (fl & 4u) phz = mat.hist[2 * hcomp + hq];
        }
        float ex = (cax * Ex[id] + cbx * (s_ex_y - s_ex_z - cw * phx)) * pecy * pecz;
        float ey = (cay * Ey[id] + cby * (s_ey_z - s_ey_x - cw * phy)) * pecz * pecx;
        const float ez = (caz * Ez[id] + cbz * (s_ez_x - s_ez_y - cw * phz)) * pecx * pecy;
        if (inj_x) ex += cbx * inj_ex;
        if (inj_y) ey += cby * inj_ey;
        Ex[id] = ex;
        Ey[id] = ey;
        Ez[id] = ez;
        hy_m = hy;
        hz_m = hz;
    }
}

// ------------------------------------------------------------------------------------------
// Point dipoles, one thread per distinct (comp, cell); its points are applied in input order
// (reference inject_dipole / inject_dipole_h, `openem/kernels/source_dft.cu:147-206`, which
// uses atomicAdd instead).
// ------------------------------------------------------------------------------------------
struct DipoleDev {
    int n_groups, nsteps1, magnetic;
    const int *g_comp, *g_cell, *g_ofs;     // per group; ofs (n_groups+1) into the point lists
    const float *coef;                      // per point
    const int *src;                         // per point
    const double *amp;                      // n_src * nsteps1
};

__global__ void __launch_bounds__(256)
k_dipoles(float *__restrict__ F0, float *__restrict__ F1, float *__restrict__ F2, DipoleDev d,
          MatDev mat, float ch, size_t ncell, const int *__restrict__ stp)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.n_groups) return;
    const int n = step_now(stp);
    if (n >= d.nsteps1) return;
    const int c = d.g_comp[t], id = d.g_cell[t];
    float *F = c == 0 ? F0 : (c == 1 ? F1 : F2);
    float scale;
    if (d.magnetic) {
        scale = ch;
    } else if (mat.mode == 0) {
        scale = mat.lut[8 * (size_t)mat.cls[id] + 2 * c + 1];
    } else {
        scale = mat.cb[(size_t)c * ncell + id];
    }
    float v = F[id];
    for (int q = d.g_ofs[t]; q < d.g_ofs[t + 1]; ++q)
        v -= scale * d.coef[q] * (float)d.amp[(size_t)d.src[q] * d.nsteps1 + n];
    F[id] = v;
}

// ------------------------------------------------------------------------------------------
// Complex-pole ADE (reference dispersion_pre / dispersion_post,
// `openem/kernels/dispersion.cu:68-147`): one thread per entry.
// ------------------------------------------------------------------------------------------
struct DispDev {
    int n_entry;
    const int *comp, *cell, *ofs, *cidx;
    const float *am1_re, *am1_im, *b_re, *b_im;
    float *p_re, *p_im;
    float *hist;
    int hb[6];
};

__device__ __forceinline__ size_t hist_index(const DispDev &d, int c, int cell, int ny, int nz) {
    const int i = cell / (ny * nz), r = cell - i * ny * nz, j = r / nz, k = r - j * nz;
    const int bx = d.hb[1] - d.hb[0], by = d.hb[3] - d.hb[2], bz = d.hb[5] - d.hb[4];
    return (((size_t)c * bx + (i - d.hb[0])) * by + (j - d.hb[2])) * bz + (k - d.hb[4]);
}

__global__ void __launch_bounds__(256)
k_disp_pre(DispDev d, const float *__restrict__ Ex, const float *__restrict__ Ey,
           const float *__restrict__ Ez, int ny, int nz)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    const int c = d.comp[e], id = d.cell[e];
    const float en = c == 0 ? Ex[id] : (c == 1 ? Ey[id] : Ez[id]);
    float w = 0.f;
    for (int t = d.ofs[e]; t < d.ofs[e + 1]; ++t) {
        const int ci = d.cidx[t];
        const float pr = d.p_re[t], pi = d.p_im[t];
        const float ar = d.am1_re[ci], ai = d.am1_im[ci];
        const float br = d.b_re[ci], bi = d.b_im[ci];
        const float dr = ar * pr - ai * pi + br * en;
        const float di = ar * pi + ai * pr + bi * en;
        d.p_re[t] = pr + dr;
        d.p_im[t] = pi + di;
        w += 2.0f * dr;
    }
    d.hist[hist_index(d, c, id, ny, nz)] = w;
}

__global__ void __launch_bounds__(256)
k_disp_post(DispDev d, const float *__restrict__ Ex, const float *__restrict__ Ey,
            const float *__restrict__ Ez)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    const int c = d.comp[e], id = d.cell[e];
    const float en = c == 0 ? Ex[id] : (c == 1 ? Ey[id] : Ez[id]);
    for (int t = d.ofs[e]; t < d.ofs[e + 1]; ++t) {
        const int ci = d.cidx[t];
        d.p_re[t] = d.p_re[t] + d.b_re[ci] * en;
        d.p_im[t] = d.p_im[t] + d.b_im[ci] * en;
    }
}

// ------------------------------------------------------------------------------------------
// Running DFT, batched: every step a snapshot of the monitored values goes into a ring of T
// float32 rows; every T steps one flush replays them in step order against the float64 phase
// table and adds into the float64 accumulators (reference accumulate_dft / accumulate_dft_box
// and the batched dft_snap_* / dft_flush_*, `openem/kernels/source_dft.cu:216-271`,
// `:371-430`, `:453-779`; there the partial sums are float32, here they are float64).
// ------------------------------------------------------------------------------------------
struct MonDev {
    int kind;                // 0 plane, 1 box
    int axis, km, n1, n2;    // plane
    int o[3], b[3];          // box
    int ncomp, npts, nf, T;
    float *snap;             // T * ncomp * npts
    double *re, *im;         // ncomp * nf * npts
    const double *phase;     // n_steps * nf * 4
};

__global__ void __launch_bounds__(256)
k_dft_snap(MonDev m, const float *__restrict__ Ex, const float *__restrict__ Ey,
           const float *__restrict__ Ez, const float *__restrict__ Hx,
           const float *__restrict__ Hy, const float *__restrict__ Hz, int ny, int nz,
           const int *__restrict__ stp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.npts) return;
    const int slot = step_now(stp) % m.T;
    float *row = m.snap + (size_t)slot * m.ncomp * m.npts;
    if (m.kind == 0) {
        const int a = q / m.n2, b = q - a * m.n2;
        int c[3];
        const int t1 = m.axis == 0 ? 1 : 0, t2 = m.axis == 2 ? 1 : 2;
        c[m.axis] = m.km; c[t1] = a; c[t2] = b;
        const size_t id = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        c[m.axis] = m.km - 1;
        const size_t idm = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        const float *Ep[3] = {Ex, Ey, Ez};
        const float *Hp[3] = {Hx, Hy, Hz};
        row[q] = Ep[t1][id];
        row[m.npts + q] = Ep[t2][id];
        row[2 * m.npts + q] = 0.5f * (Hp[t1][id] + Hp[t1][idm]);
        row[3 * m.npts + q] = 0.5f * (Hp[t2][id] + Hp[t2][idm]);
    } else {
        const int bjk = m.b[1] * m.b[2];
        const int a = q / bjk, r = q - a * bjk, b = r / m.b[2], cc = r - b * m.b[2];
        const size_t id = ((size_t)(m.o[0] + a) * ny + (m.o[1] + b)) * nz + (m.o[2] + cc);
        row[q] = Ex[id];
        row[m.npts + q] = Ey[id];
        row[2 * (size_t)m.npts + q] = Ez[id];
        row[3 * (size_t)m.npts + q] = Hx[id];
        row[4 * (size_t)m.npts + q] = Hy[id];
        row[5 * (size_t)m.npts + q] = Hz[id];
    }
}

// grid: x over points, y over frequencies.  Launched after the step counter advanced past the
// last snapshot: the cnt snapshots ending at step *stp - 1 are replayed in step order; NC
// components of which the first NC/2 are E (whole-step phase).
template <int NC>
__global__ void __launch_bounds__(256)
k_dft_flush(MonDev m, int cnt, const int *__restrict__ stp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (q >= m.npts || f >= m.nf) return;
    const int last = step_now(stp) - 1;
    const int first = last - cnt + 1;
    double sre[NC], sim[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) { sre[c] = 0.0; sim[c] = 0.0; }
    for (int s = first; s <= last; ++s) {
        const float *row = m.snap + (size_t)(s % m.T) * NC * m.npts;
        const double *ph = m.phase + ((size_t)s * m.nf + f) * 4;
        const double ce = ph[0], se = ph[1], chh = ph[2], shh = ph[3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double v = (double)row[(size_t)c * m.npts + q];
            sre[c] += v * (c < NC / 2 ? ce : chh);
            sim[c] += v * (c < NC / 2 ? se : shh);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const size_t o = ((size_t)c * m.nf + f) * m.npts + q;
        m.re[o] += sre[c];
        m.im[o] += sim[c];
    }
}

__global__ void k_step_advance(int *stp) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *stp += 1;
}
