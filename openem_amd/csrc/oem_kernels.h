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

#define OEM_NXCD 8
#define OEM_MAX_PW 2

// Per-axis tables interleaved as tab[i*AX_W + field]: one pointer per axis keeps the kernel
// argument block (and the SGPRs holding it) small.
enum { AX_IDLP = 0, AX_IDLD, AX_MNX, AX_MPV, AX_PEC, AX_AE, AX_BE, AX_KE, AX_AH, AX_BH, AX_KH,
       AX_W = 12 };

struct AxisDev {
    const float *tab;       // n * AX_W
    const int *nxt, *prv;
    int n, lo, hi;
};

__device__ __forceinline__ float axv(const AxisDev &A, int i, int f) { return A.tab[i * AX_W + f]; }

struct GridDev {
    AxisDev ax[3];
    int nx, ny, nz;
    int plane;          // ny*nz
    int tiles;          // k-tiles * j-tiles of the (ny, nz) plane
    int chunk;          // i-planes per block
    int nblocks;        // tiles * ceil(nx / chunk)
};

// psi slabs: one allocation per side (H, E) holding six sub-arrays at element offsets
// off[2*axis + c], where c selects the field component pair member:
//   H side: x -> (Hy, Hz), y -> (Hx, Hz), z -> (Hx, Hy); E side likewise with E.
// x slab (wx, ny, nz), y slab (nx, wy, nz), z slab (nx, ny, wz), w = lo + hi layers.
struct PsiDev {
    float *h, *e;
    uint32_t off[6];
    uint32_t total;     // elements per side
};

struct PlaneSrcDev {
    int n;              // active plane sources (0..OEM_MAX_PW)
    int pol[OEM_MAX_PW], kh[OEM_MAX_PW], ks[OEM_MAX_PW];
    float coef_h[OEM_MAX_PW], coef_e[OEM_MAX_PW];
    const double *ex_inc[OEM_MAX_PW], *hy_inc[OEM_MAX_PW];
};

struct MatDev {
    int mode;           // 0: class LUT, 1: dense ca/cb
    int n_classes;
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

// Buffer resources: an SGPR descriptor + a fixed 32-bit per-lane VGPR offset + a per-iteration
// SGPR offset.  This keeps the x march's address arithmetic out of the VGPRs (the compiler
// would otherwise carry one 64-bit pointer per stream).  gfx9-family descriptor word 3.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
#define OEM_RSRC_W3 0x00020000

__device__ __forceinline__ rsrc_t mk_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             OEM_RSRC_W3);
}
__device__ __forceinline__ float bld(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ uint32_t bld_u16(rsrc_t r, uint32_t voff, uint32_t soff) {
    return (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
}
__device__ __forceinline__ void bst(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, 0);
}

// Tile of the (ny, nz) plane owned by one block: OEM_TK k-lanes x OEM_TJ j-rows.
#define OEM_TK 64
#define OEM_TJ 4
#define OEM_LDS_CLASSES 512

struct TileCoord {
    int j, k, i0, i1;
    bool ok;
};

__device__ __forceinline__ TileCoord tile_coord(const GridDev &g) {
    const int lb = xcd_remap(blockIdx.x, g.nblocks);
    const int nkt = (g.nz + OEM_TK - 1) / OEM_TK;
    const int tile = lb % g.tiles, chunk = lb / g.tiles;
    const int jt = tile / nkt, kt = tile - jt * nkt;
    TileCoord t;
    t.k = kt * OEM_TK + (int)threadIdx.x;
    t.j = jt * OEM_TJ + (int)threadIdx.y;
    t.i0 = chunk * g.chunk;
    t.i1 = min(t.i0 + g.chunk, g.nx);
    t.ok = t.k < g.nz && t.j < g.ny;
    return t;
}

// CPML recursion on one psi element already loaded: psi' = b*psi + a*d, s = invk*d + psi'.
__device__ __forceinline__ float stretch_v(float &psi, float d, float a, float b, float invk) {
    psi = b * psi + a * d;
    return invk * d + psi;
}

// ------------------------------------------------------------------------------------------
// H half step (reference update_h, `openem/kernels/yee.cu:76-137`) fused with the H side of
// the plane-wave TF/SF correction (`openem/kernels/source_dft.cu:40-55`).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_update_h(float *__restrict__ H, const float *__restrict__ E, PsiDev ps, GridDev g, float ch,
           PlaneSrcDev pw, const int *__restrict__ stp)
{
    const TileCoord tc = tile_coord(g);
    if (!tc.ok) return;
    const int j = tc.j, k = tc.k;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const size_t ncell = (size_t)g.plane * nx;
    const uint32_t cb = (uint32_t)(ncell * 4);
    const rsrc_t rEx = mk_rsrc(E, cb), rEy = mk_rsrc(E + ncell, cb), rEz = mk_rsrc(E + 2 * ncell, cb);
    const rsrc_t rHx = mk_rsrc(H, cb), rHy = mk_rsrc(H + ncell, cb), rHz = mk_rsrc(H + 2 * ncell, cb);
    const rsrc_t rP = mk_rsrc(ps.h, ps.total * 4u);

    const uint32_t op = (uint32_t)(j * nz + k) * 4u;              // own cell in the plane
    const uint32_t oj = (uint32_t)(Y.nxt[j] * nz + k) * 4u;       // (j+1, k)
    const uint32_t ok = (uint32_t)(j * nz + Z.nxt[k]) * 4u;       // (j, k+1)
    const float mj = axv(Y, j, AX_MNX), mk = axv(Z, k, AX_MNX);
    const float idy = axv(Y, j, AX_IDLP), idz = axv(Z, k, AX_IDLP);
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = axv(Y, j, AX_AH); by = axv(Y, j, AX_BH); ky = axv(Y, j, AX_KH); }
    if (sz >= 0) { az = axv(Z, k, AX_AH); bz = axv(Z, k, AX_BH); kz = axv(Z, k, AX_KH); }
    const uint32_t oy = (uint32_t)(max(sy, 0) * nz + k) * 4u;     // psi y-slab, within plane i
    const uint32_t oz = (uint32_t)(j * wz + max(sz, 0)) * 4u;     // psi z-slab, within plane i

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

    const uint32_t plane4 = (uint32_t)g.plane * 4u;
    float ey = bld(rEy, op, (uint32_t)tc.i0 * plane4);
    float ez = bld(rEz, op, (uint32_t)tc.i0 * plane4);
    for (int i = tc.i0; i < tc.i1; ++i) {
        const uint32_t so = (uint32_t)i * plane4;
        const uint32_t sn = (uint32_t)X.nxt[i] * plane4;
        const float ex = bld(rEx, op, so), ex_j = bld(rEx, oj, so), ex_k = bld(rEx, ok, so);
        const float ey_k = bld(rEy, ok, so), ez_j = bld(rEz, oj, so);
        const float ey_n = bld(rEy, op, sn), ez_n = bld(rEz, op, sn);
        float hx = bld(rHx, op, so), hy = bld(rHy, op, so), hz = bld(rHz, op, so);
        const int sx = slab_of(i, nx, X.lo, X.hi);
        const uint32_t sx0 = (ps.off[0] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * 4u;
        const uint32_t sx1 = (ps.off[1] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * 4u;
        const uint32_t sy0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * nz)) * 4u;
        const uint32_t sy1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * nz)) * 4u;
        const uint32_t sz0 = (ps.off[4] + (uint32_t)i * (uint32_t)(ny * wz)) * 4u;
        const uint32_t sz1 = (ps.off[5] + (uint32_t)i * (uint32_t)(ny * wz)) * 4u;
        float psx0 = 0.f, psx1 = 0.f, psy0 = 0.f, psy1 = 0.f, psz0 = 0.f, psz1 = 0.f;
        if (sx >= 0) { psx0 = bld(rP, op, sx0); psx1 = bld(rP, op, sx1); }
        if (sy >= 0) { psy0 = bld(rP, oy, sy0); psy1 = bld(rP, oy, sy1); }
        if (sz >= 0) { psz0 = bld(rP, oz, sz0); psz1 = bld(rP, oz, sz1); }

        const float mi = axv(X, i, AX_MNX), idx = axv(X, i, AX_IDLP);
        const float dzy = (ez_j * mj - ez) * idy;
        const float dyz = (ey_k * mk - ey) * idz;
        const float dxz = (ex_k * mk - ex) * idz;
        const float dzx = (ez_n * mi - ez) * idx;
        const float dyx = (ey_n * mi - ey) * idx;
        const float dxy = (ex_j * mj - ex) * idy;

        float s_hx_y = dzy, s_hz_y = dxy, s_hx_z = dyz, s_hy_z = dxz, s_hy_x = dzx, s_hz_x = dyx;
        if (sx >= 0) {
            const float ax_ = axv(X, i, AX_AH), bx_ = axv(X, i, AX_BH), kx_ = axv(X, i, AX_KH);
            s_hy_x = stretch_v(psx0, dzx, ax_, bx_, kx_);
            s_hz_x = stretch_v(psx1, dyx, ax_, bx_, kx_);
            bst(rP, psx0, op, sx0);
            bst(rP, psx1, op, sx1);
        }
        if (sy >= 0) {
            s_hx_y = stretch_v(psy0, dzy, ay, by, ky);
            s_hz_y = stretch_v(psy1, dxy, ay, by, ky);
            bst(rP, psy0, oy, sy0);
            bst(rP, psy1, oy, sy1);
        }
        if (sz >= 0) {
            s_hx_z = stretch_v(psz0, dyz, az, bz, kz);
            s_hy_z = stretch_v(psz1, dxz, az, bz, kz);
            bst(rP, psz0, oz, sz0);
            bst(rP, psz1, oz, sz1);
        }
        hx = hx - ch * (s_hx_y - s_hx_z) + inj_hx;
        hy = hy - ch * (s_hy_z - s_hy_x) + inj_hy;
        hz = hz - ch * (s_hz_x - s_hz_y);
        bst(rHx, hx, op, so);
        bst(rHy, hy, op, so);
        bst(rHz, hz, op, so);
        ey = ey_n;
        ez = ez_n;
    }
}

// ------------------------------------------------------------------------------------------
// E whole step (reference update_e / update_e_lut, `openem/kernels/yee.cu:154-337`) fused with
// the E side of the plane-wave TF/SF correction (`openem/kernels/source_dft.cu:58-73`).  The
// class LUT is staged in LDS and the next plane's class index is prefetched, so the per-cell
// coefficient gather never waits on a global load issued in the same iteration.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_update_e(float *__restrict__ E, const float *__restrict__ H, PsiDev ps, GridDev g,
           MatDev mat, float cw, PlaneSrcDev pw, const int *__restrict__ stp)
{
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    const bool lds_lut = mat.mode == 0 && mat.n_classes <= OEM_LDS_CLASSES;
    if (lds_lut) {
        const int tid = threadIdx.y * OEM_TK + threadIdx.x;
        const float4 *src = reinterpret_cast<const float4 *>(mat.lut);
        for (int q = tid; q < 2 * mat.n_classes; q += OEM_TK * OEM_TJ) s_lut[q] = src[q];
        __syncthreads();
    }
    const TileCoord tc = tile_coord(g);
    if (!tc.ok) return;
    const int j = tc.j, k = tc.k;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const size_t ncell = (size_t)g.plane * nx;
    const uint32_t cbytes = (uint32_t)(ncell * 4);
    const rsrc_t rEx = mk_rsrc(E, cbytes), rEy = mk_rsrc(E + ncell, cbytes),
                 rEz = mk_rsrc(E + 2 * ncell, cbytes);
    const rsrc_t rHx = mk_rsrc(H, cbytes), rHy = mk_rsrc(H + ncell, cbytes),
                 rHz = mk_rsrc(H + 2 * ncell, cbytes);
    const rsrc_t rP = mk_rsrc(ps.e, ps.total * 4u);
    const rsrc_t rC = mk_rsrc(mat.cls, (uint32_t)(ncell * 2));

    const uint32_t op = (uint32_t)(j * nz + k) * 4u;
    const uint32_t oj = (uint32_t)(Y.prv[j] * nz + k) * 4u;       // (j-1, k)
    const uint32_t ok = (uint32_t)(j * nz + Z.prv[k]) * 4u;       // (j, k-1)
    const float mj = axv(Y, j, AX_MPV), mk = axv(Z, k, AX_MPV);
    const float idy = axv(Y, j, AX_IDLD), idz = axv(Z, k, AX_IDLD);
    const float pecy = axv(Y, j, AX_PEC), pecz = axv(Z, k, AX_PEC);
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = axv(Y, j, AX_AE); by = axv(Y, j, AX_BE); ky = axv(Y, j, AX_KE); }
    if (sz >= 0) { az = axv(Z, k, AX_AE); bz = axv(Z, k, AX_BE); kz = axv(Z, k, AX_KE); }
    const uint32_t oy = (uint32_t)(max(sy, 0) * nz + k) * 4u;
    const uint32_t oz = (uint32_t)(j * wz + max(sz, 0)) * 4u;

    float inj_ex = 0.f, inj_ey = 0.f;
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
    const int hbx = mat.hb[1] - mat.hb[0], hby = mat.hb[3] - mat.hb[2],
              hbz = mat.hb[5] - mat.hb[4];
    const bool hrow = mat.hist && j >= mat.hb[2] && j < mat.hb[3] && k >= mat.hb[4] &&
                      k < mat.hb[5];
    const uint32_t hcomp = (uint32_t)(hbx * hby * hbz);
    const rsrc_t rHist = mk_rsrc(mat.hist, 3u * hcomp * 4u);
    const uint32_t oh = hrow ? (uint32_t)((j - mat.hb[2]) * hbz + (k - mat.hb[4])) * 4u : 0u;
    const uint32_t oc = (uint32_t)(j * nz + k) * 2u;               // uint16 class index

    const uint32_t plane4 = (uint32_t)g.plane * 4u, plane2 = (uint32_t)g.plane * 2u;
    const uint32_t sp0 = (uint32_t)X.prv[tc.i0] * plane4;
    float hy_m = bld(rHy, op, sp0), hz_m = bld(rHz, op, sp0);
    uint32_t cls_next = mat.mode == 0 ? bld_u16(rC, oc, (uint32_t)tc.i0 * plane2) : 0u;
    for (int i = tc.i0; i < tc.i1; ++i) {
        const uint32_t so = (uint32_t)i * plane4;
        const uint32_t cls = cls_next;
        if (mat.mode == 0 && i + 1 < tc.i1) cls_next = bld_u16(rC, oc, so / 2u + plane2);
        const float hx = bld(rHx, op, so), hy = bld(rHy, op, so), hz = bld(rHz, op, so);
        const float hx_j = bld(rHx, oj, so), hx_k = bld(rHx, ok, so);
        const float hy_k = bld(rHy, ok, so), hz_j = bld(rHz, oj, so);
        float ex = bld(rEx, op, so), ey = bld(rEy, op, so), ez = bld(rEz, op, so);
        const int sx = slab_of(i, nx, X.lo, X.hi);
        const uint32_t sx0 = (ps.off[0] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * 4u;
        const uint32_t sx1 = (ps.off[1] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * 4u;
        const uint32_t sy0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * nz)) * 4u;
        const uint32_t sy1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * nz)) * 4u;
        const uint32_t sz0 = (ps.off[4] + (uint32_t)i * (uint32_t)(ny * wz)) * 4u;
        const uint32_t sz1 = (ps.off[5] + (uint32_t)i * (uint32_t)(ny * wz)) * 4u;
        float psx0 = 0.f, psx1 = 0.f, psy0 = 0.f, psy1 = 0.f, psz0 = 0.f, psz1 = 0.f;
        if (sx >= 0) { psx0 = bld(rP, op, sx0); psx1 = bld(rP, op, sx1); }
        if (sy >= 0) { psy0 = bld(rP, oy, sy0); psy1 = bld(rP, oy, sy1); }
        if (sz >= 0) { psz0 = bld(rP, oz, sz0); psz1 = bld(rP, oz, sz1); }

        float cax, cbx, cay, cby, caz, cbz;
        uint32_t fl;
        if (lds_lut) {
            const float4 r0 = s_lut[2 * cls], r1 = s_lut[2 * cls + 1];
            cax = r0.x; cbx = r0.y; cay = r0.z; cby = r0.w; caz = r1.x; cbz = r1.y;
            fl = __float_as_uint(r1.z);
        } else if (mat.mode == 0) {
            const float4 *row = reinterpret_cast<const float4 *>(mat.lut + 8u * cls);
            const float4 r0 = row[0], r1 = row[1];
            cax = r0.x; cbx = r0.y; cay = r0.z; cby = r0.w; caz = r1.x; cbz = r1.y;
            fl = __float_as_uint(r1.z);
        } else {
            const size_t id = (size_t)i * g.plane + (size_t)j * nz + k;
            cax = mat.ca[id]; cay = mat.ca[ncell + id]; caz = mat.ca[2 * ncell + id];
            cbx = mat.cb[id]; cby = mat.cb[ncell + id]; cbz = mat.cb[2 * ncell + id];
            fl = (hrow && i >= mat.hb[0] && i < mat.hb[1]) ? 7u : 0u;
        }
        float phx = 0.f, phy = 0.f, phz = 0.f;
        if (fl) {
            const uint32_t sh = (uint32_t)(i - mat.hb[0]) * (uint32_t)(hby * hbz) * 4u;
            if (fl & 1u) phx = bld(rHist, oh, sh);
            if (fl & 2u) phy = bld(rHist, oh, sh + hcomp * 4u);
            if (fl & 4u) phz = bld(rHist, oh, sh + 2u * hcomp * 4u);
        }

        const float mi = axv(X, i, AX_MPV), idx = axv(X, i, AX_IDLD), pecx = axv(X, i, AX_PEC);
        const float dzy = (hz - hz_j * mj) * idy;
        const float dyz = (hy - hy_k * mk) * idz;
        const float dxz = (hx - hx_k * mk) * idz;
        const float dzx = (hz - hz_m * mi) * idx;
        const float dyx = (hy - hy_m * mi) * idx;
        const float dxy = (hx - hx_j * mj) * idy;

        float s_ex_y = dzy, s_ez_y = dxy, s_ex_z = dyz, s_ey_z = dxz, s_ey_x = dzx, s_ez_x = dyx;
        if (sx >= 0) {
            const float ax_ = axv(X, i, AX_AE), bx_ = axv(X, i, AX_BE), kx_ = axv(X, i, AX_KE);
            s_ey_x = stretch_v(psx0, dzx, ax_, bx_, kx_);
            s_ez_x = stretch_v(psx1, dyx, ax_, bx_, kx_);
            bst(rP, psx0, op, sx0);
            bst(rP, psx1, op, sx1);
        }
        if (sy >= 0) {
            s_ex_y = stretch_v(psy0, dzy, ay, by, ky);
            s_ez_y = stretch_v(psy1, dxy, ay, by, ky);
            bst(rP, psy0, oy, sy0);
            bst(rP, psy1, oy, sy1);
        }
        if (sz >= 0) {
            s_ex_z = stretch_v(psz0, dyz, az, bz, kz);
            s_ey_z = stretch_v(psz1, dxz, az, bz, kz);
            bst(rP, psz0, oz, sz0);
            bst(rP, psz1, oz, sz1);
        }
        ex = (cax * ex + cbx * (s_ex_y - s_ex_z - cw * phx)) * pecy * pecz;
        ey = (cay * ey + cby * (s_ey_z - s_ey_x - cw * phy)) * pecz * pecx;
        ez = (caz * ez + cbz * (s_ez_x - s_ez_y - cw * phz)) * pecx * pecy;
        if (inj_x) ex += cbx * inj_ex;
        if (inj_y) ey += cby * inj_ey;
        bst(rEx, ex, op, so);
        bst(rEy, ey, op, so);
        bst(rEz, ez, op, so);
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

// post(n) + pre(n+1) in one pass (reference dispersion_step, `openem/kernels/dispersion.cu:168-207`):
// nothing writes E between the two, so the same E^{n+1} feeds both and P is read and written
// once instead of twice.
__global__ void __launch_bounds__(256)
k_disp_step(DispDev d, const float *__restrict__ Ex, const float *__restrict__ Ey,
            const float *__restrict__ Ez, int ny, int nz)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    const int c = d.comp[e], id = d.cell[e];
    const float en = c == 0 ? Ex[id] : (c == 1 ? Ey[id] : Ez[id]);
    float w = 0.f;
    for (int t = d.ofs[e]; t < d.ofs[e + 1]; ++t) {
        const int ci = d.cidx[t];
        const float br = d.b_re[ci], bi = d.b_im[ci];
        const float pr = d.p_re[t] + br * en;
        const float pi = d.p_im[t] + bi * en;
        const float ar = d.am1_re[ci], ai = d.am1_im[ci];
        const float dr = ar * pr - ai * pi + br * en;
        const float di = ar * pi + ai * pr + bi * en;
        d.p_re[t] = pr + dr;
        d.p_im[t] = pi + di;
        w += 2.0f * dr;
    }
    d.hist[hist_index(d, c, id, ny, nz)] = w;
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
