// Device kernels of the MI355X (gfx950, CDNA4) FDTD stepper.
//
// The Yee update is HBM-bound (about 1 flop per byte), so the design goal is: every field
// byte crosses HBM once per sweep.  Each work-group owns a 64 (k) x 8 (j) tile of the (ny, nz)
// plane and marches along x (the slowest axis) over a chunk of i-planes.  The x neighbour plane
// the curl needs is carried in registers from one iteration to the next; the y/z neighbours are
// in the same or an adjacent tile and are served by L1/L2.  The linear block id is remapped so
// that consecutive tiles land on the same XCD (MI355X dispatches work-groups round-robin over 8
// XCDs, each with its own L2), which keeps the halo rows of neighbouring tiles in one L2.
//
// Every kernel is a template on the field type F: float for real scenes, cf2 (re, im
// interleaved) for Bloch scenes.  All coefficients are real; the Bloch phase enters only through
// the complex masks at the periodic wrap (the reference's layout of the complex path,
// `openem/kernels/bloch.cu:1-23`).
//
// CPML psi is stored only inside the PML slabs of each axis.  Outside a slab the reference's
// identity coefficients (a=0, b=1, 1/kappa=1, `openem/kernels/yee.cu:60-74`) make the stretched
// derivative equal to the plain one, which is what the kernels compute there.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OEM_NXCD 8
#define OEM_MAX_PW 2
#define OEM_TK 64
#ifndef OEM_TJ
#define OEM_TJ 8
#endif
#define OEM_LDS_CLASSES 512

// ---------------------------------------------------------------- complex field element ----
struct cf2 {
    float r, i;
};
__device__ __forceinline__ cf2 operator+(cf2 a, cf2 b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cf2 operator-(cf2 a, cf2 b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cf2 operator*(float s, cf2 a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ cf2 operator*(cf2 a, float s) { return {a.r * s, a.i * s}; }
__device__ __forceinline__ cf2 operator-(cf2 a) { return {-a.r, -a.i}; }
__device__ __forceinline__ cf2 &operator+=(cf2 &a, cf2 b) { a.r += b.r; a.i += b.i; return a; }
__device__ __forceinline__ cf2 &operator-=(cf2 &a, cf2 b) { a.r -= b.r; a.i -= b.i; return a; }

// ---------------------------------------------------------------- buffer resources ----------
// An SGPR descriptor + a fixed 32-bit per-lane VGPR offset + a per-iteration SGPR offset keeps
// the x march's address arithmetic out of the VGPRs.  gfx9-family descriptor word 3.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
#define OEM_RSRC_W3 0x00020000

__device__ __forceinline__ rsrc_t mk_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             OEM_RSRC_W3);
}
__device__ __forceinline__ uint32_t bld_u16(rsrc_t r, uint32_t voff, uint32_t soff) {
    return (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
}

// The 64-bit buffer builtins take and return a GCC vector of two uint32; converting that
// implicitly to or from an ext_vector type splats one lane, so every conversion goes through
// __builtin_bit_cast.
typedef unsigned int oem_u32x2 __attribute__((vector_size(8)));

// Cache policy of every buffer store: non-temporal (cache-policy bit 1).  Nothing a kernel
// stores is read again in the same launch, so the stored lines need not displace, in L2, the
// lines that neighbouring work-groups read again (the fused march's halo and neighbour lines):
// on the 512^3 step its reads fell from 2.29 to 2.17 M KiB per launch and the march from 1.729
// to 1.703 ms; the Bloch step from 1.105 to 1.053 ms.
#ifndef OEM_ST_AUX
#define OEM_ST_AUX 2
#endif

template <typename F> struct FT;
template <> struct FT<float> {
    static constexpr uint32_t B = 4;
    static __device__ __forceinline__ float ld(rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
    }
    static __device__ __forceinline__ void st(rsrc_t r, float x, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, x), r, v, s, OEM_ST_AUX);
    }
    static __device__ __forceinline__ float mmul(float x, float mr, float) { return x * mr; }
    static __device__ __forceinline__ float zero() { return 0.f; }
    static __device__ __forceinline__ float addre(float x, float a) { return x + a; }
    // x + s * Re(v) (real fields keep the real part of a complex correction)
    static __device__ __forceinline__ float addc(float x, float s, float vr, float) {
        return x + s * vr;
    }
};
template <> struct FT<cf2> {
    static constexpr uint32_t B = 8;
    static __device__ __forceinline__ cf2 ld(rsrc_t r, uint32_t v, uint32_t s) {
        const oem_u32x2 q = __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0);
        return __builtin_bit_cast(cf2, q);
    }
    static __device__ __forceinline__ void st(rsrc_t r, cf2 x, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(oem_u32x2, x), r, v, s, OEM_ST_AUX);
    }
    static __device__ __forceinline__ cf2 mmul(cf2 x, float mr, float mi) {
        return {x.r * mr - x.i * mi, x.r * mi + x.i * mr};
    }
    static __device__ __forceinline__ cf2 zero() { return {0.f, 0.f}; }
    static __device__ __forceinline__ cf2 addre(cf2 x, float a) { return {x.r + a, x.i}; }
    static __device__ __forceinline__ cf2 addc(cf2 x, float s, float vr, float vi) {
        return {x.r + s * vr, x.i + s * vi};
    }
};

// ---------------------------------------------------------------- tables -------------------
// Per-axis tables interleaved as tab[i*AX_W + field]: one pointer per axis keeps the kernel
// argument block (and the SGPRs holding it) small.  MNXI/MPVI: imaginary parts of the masks
// (non-zero only at a Bloch wrap).
enum { AX_IDLP = 0, AX_IDLD, AX_MNX, AX_MPV, AX_PEC, AX_AE, AX_BE, AX_KE, AX_AH, AX_BH, AX_KH,
       AX_MNXI, AX_MPVI, AX_W = 16 };

struct AxisDev {
    const float *tab;       // n * AX_W
    const int *nxt, *prv;
    int n, lo, hi;
};

__device__ __forceinline__ float axv(const AxisDev &A, int i, int f) { return A.tab[i * AX_W + f]; }

// A table row read through the constant address space: the tables never change during a
// launch, so a wave-uniform row comes in through scalar loads, whose wait does not hold the
// vector loads and stores in flight (those share one in-order counter).
typedef const __attribute__((address_space(4))) float *ctab_t;
__device__ __forceinline__ ctab_t ctab_row(const AxisDev &A, int i) {
    return (ctab_t)(A.tab + (size_t)i * AX_W);
}
// A wave-uniform entry of a read-only index table, through a scalar load.
__device__ __forceinline__ int cidx_at(const int *p, int i) {
    return ((const __attribute__((address_space(4))) int *)p)[i];
}

struct GridDev {
    AxisDev ax[3];
    int nx, ny, nz;
    int plane;          // ny*nz
    int tiles;          // k-tiles * j-tiles of the (ny, nz) plane (flat: blocks of the plane)
    int chunk;          // i-planes per block
    int nblocks;        // tiles * ceil(nx / chunk)
    int flat;           // the sweeps number the (j, k) columns along the plane (nz % 64 != 0)
};

// psi slabs: one allocation per side (H, E) holding six sub-arrays at element offsets
// off[2*axis + c], where c selects the field component pair member:
//   H side: x -> (Hy, Hz), y -> (Hx, Hz), z -> (Hx, Hy); E side likewise with E.
// x slab (wx, ny, nz), y slab (nx, wy, nz), z slab (nx, ny, wz), w = lo + hi layers.
struct PsiDev {
    void *h, *e;
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
    void *hist;         // (3, bx, by, bz) elements of F inside the hist box, or nullptr
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

__device__ __forceinline__ int step_now(const int *stp) { return *stp; }

// Next index along an axis: interior neighbours are i + 1 (the host checks), only the last
// cell's neighbour comes from the table (`last` = A.nxt[n - 1], loaded once and marked
// wave-uniform).  A table load inside the x march would make the next plane's offset a
// per-lane value, and every load offset by it a per-lane loop behind a full memory wait.
__device__ __forceinline__ int next_of(int i, int n, int last) { return i + 1 < n ? i + 1 : last; }

struct TileCoord {
    int j, k, i0, i1;
    bool ok;
};

// Logical block lb of a scene -> (j, k) of this thread and the i-range of the x march: an
// OEM_TJ x OEM_TK tile of the (j, k) plane, or, when nz is not a multiple of OEM_TK (a 48-cell
// slab would leave a quarter of every wave idle), the block's OEM_TJ * OEM_TK consecutive
// columns of the plane.
__device__ __forceinline__ TileCoord tile_coord(const GridDev &g, int lb) {
    const int tile = lb % g.tiles, chunk = lb / g.tiles;
    TileCoord t;
    if (g.flat) {
        const int q = tile * (OEM_TK * OEM_TJ) + (int)threadIdx.y * OEM_TK + (int)threadIdx.x;
        t.ok = q < g.plane;
        t.j = t.ok ? q / g.nz : 0;
        t.k = t.ok ? q - t.j * g.nz : 0;
    } else {
        const int nkt = (g.nz + OEM_TK - 1) / OEM_TK;
        const int jt = tile / nkt, kt = tile - jt * nkt;
        t.k = kt * OEM_TK + (int)threadIdx.x;
        t.j = jt * OEM_TJ + (int)threadIdx.y;
        t.ok = t.k < g.nz && t.j < g.ny;
    }
    t.i0 = chunk * g.chunk;
    t.i1 = min(t.i0 + g.chunk, g.nx);
    return t;
}

// CPML recursion on one psi element already loaded: psi' = b*psi + a*d, s = invk*d + psi'.
template <typename F>
__device__ __forceinline__ F stretch_v(F &psi, F d, float a, float b, float invk) {
    psi = b * psi + a * d;
    return invk * d + psi;
}

// ------------------------------------------------------------------------------------------
// H half step (reference update_h, `openem/kernels/yee.cu:76-137`; complex variant
// update_h_c, `openem/kernels/bloch.cu:38-126`) fused with the H side of the plane-wave TF/SF
// correction (`openem/kernels/source_dft.cu:40-55`).
// ------------------------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ void update_h_tile(F *__restrict__ H, const F *__restrict__ E,
                                              const PsiDev &ps, const GridDev &g, float ch,
                                              const PlaneSrcDev &pw, const int *__restrict__ stp,
                                              const TileCoord &tc)
{
    typedef FT<F> T;
    const int j = tc.j, k = tc.k;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const size_t ncell = (size_t)g.plane * nx;
    const uint32_t cb = (uint32_t)(ncell * T::B);
    const rsrc_t rEx = mk_rsrc(E, cb), rEy = mk_rsrc(E + ncell, cb), rEz = mk_rsrc(E + 2 * ncell, cb);
    const rsrc_t rHx = mk_rsrc(H, cb), rHy = mk_rsrc(H + ncell, cb), rHz = mk_rsrc(H + 2 * ncell, cb);
    const rsrc_t rP = mk_rsrc(ps.h, ps.total * T::B);

    const uint32_t op = (uint32_t)(j * nz + k) * T::B;            // own cell in the plane
    const uint32_t oj = (uint32_t)(Y.nxt[j] * nz + k) * T::B;     // (j+1, k)
    const uint32_t ok = (uint32_t)(j * nz + Z.nxt[k]) * T::B;     // (j, k+1)
    const float mj = axv(Y, j, AX_MNX), mk = axv(Z, k, AX_MNX);
    const float mji = axv(Y, j, AX_MNXI), mki = axv(Z, k, AX_MNXI);
    const float idy = axv(Y, j, AX_IDLP), idz = axv(Z, k, AX_IDLP);
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = axv(Y, j, AX_AH); by = axv(Y, j, AX_BH); ky = axv(Y, j, AX_KH); }
    if (sz >= 0) { az = axv(Z, k, AX_AH); bz = axv(Z, k, AX_BH); kz = axv(Z, k, AX_KH); }
    const uint32_t oy = (uint32_t)(max(sy, 0) * nz + k) * T::B;   // psi y-slab, within plane i
    const uint32_t oz = (uint32_t)(j * wz + max(sz, 0)) * T::B;   // psi z-slab, within plane i

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

    const uint32_t planeB = (uint32_t)g.plane * T::B;
    const int xn_last = __builtin_amdgcn_readfirstlane(X.nxt[nx - 1]);
    F ey = T::ld(rEy, op, (uint32_t)tc.i0 * planeB);
    F ez = T::ld(rEz, op, (uint32_t)tc.i0 * planeB);
    for (int i = tc.i0; i < tc.i1; ++i) {
        const uint32_t so = (uint32_t)i * planeB;
        const uint32_t sn = (uint32_t)next_of(i, nx, xn_last) * planeB;
        const F ex = T::ld(rEx, op, so), ex_j = T::ld(rEx, oj, so), ex_k = T::ld(rEx, ok, so);
        const F ey_k = T::ld(rEy, ok, so), ez_j = T::ld(rEz, oj, so);
        const F ey_n = T::ld(rEy, op, sn), ez_n = T::ld(rEz, op, sn);
        F hx = T::ld(rHx, op, so), hy = T::ld(rHy, op, so), hz = T::ld(rHz, op, so);
        const int sx = slab_of(i, nx, X.lo, X.hi);
        const uint32_t sx0 = (ps.off[0] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sx1 = (ps.off[1] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sy0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sy1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sz0 = (ps.off[4] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        const uint32_t sz1 = (ps.off[5] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        F psx0 = T::zero(), psx1 = T::zero(), psy0 = T::zero(), psy1 = T::zero(),
          psz0 = T::zero(), psz1 = T::zero();
        if (sx >= 0) { psx0 = T::ld(rP, op, sx0); psx1 = T::ld(rP, op, sx1); }
        if (sy >= 0) { psy0 = T::ld(rP, oy, sy0); psy1 = T::ld(rP, oy, sy1); }
        if (sz >= 0) { psz0 = T::ld(rP, oz, sz0); psz1 = T::ld(rP, oz, sz1); }

        const ctab_t xr = ctab_row(X, i);
        const float mi = xr[AX_MNX], mii = xr[AX_MNXI], idx = xr[AX_IDLP];
        const F dzy = (T::mmul(ez_j, mj, mji) - ez) * idy;
        const F dyz = (T::mmul(ey_k, mk, mki) - ey) * idz;
        const F dxz = (T::mmul(ex_k, mk, mki) - ex) * idz;
        const F dzx = (T::mmul(ez_n, mi, mii) - ez) * idx;
        const F dyx = (T::mmul(ey_n, mi, mii) - ey) * idx;
        const F dxy = (T::mmul(ex_j, mj, mji) - ex) * idy;

        F s_hx_y = dzy, s_hz_y = dxy, s_hx_z = dyz, s_hy_z = dxz, s_hy_x = dzx, s_hz_x = dyx;
        if (sx >= 0) {
            const float ax_ = xr[AX_AH], bx_ = xr[AX_BH], kx_ = xr[AX_KH];
            s_hy_x = stretch_v(psx0, dzx, ax_, bx_, kx_);
            s_hz_x = stretch_v(psx1, dyx, ax_, bx_, kx_);
            T::st(rP, psx0, op, sx0);
            T::st(rP, psx1, op, sx1);
        }
        if (sy >= 0) {
            s_hx_y = stretch_v(psy0, dzy, ay, by, ky);
            s_hz_y = stretch_v(psy1, dxy, ay, by, ky);
            T::st(rP, psy0, oy, sy0);
            T::st(rP, psy1, oy, sy1);
        }
        if (sz >= 0) {
            s_hx_z = stretch_v(psz0, dyz, az, bz, kz);
            s_hy_z = stretch_v(psz1, dxz, az, bz, kz);
            T::st(rP, psz0, oz, sz0);
            T::st(rP, psz1, oz, sz1);
        }
        hx = T::addre(hx - ch * (s_hx_y - s_hx_z), inj_hx);
        hy = T::addre(hy - ch * (s_hy_z - s_hy_x), inj_hy);
        hz = hz - ch * (s_hz_x - s_hz_y);
        T::st(rHx, hx, op, so);
        T::st(rHy, hy, op, so);
        T::st(rHz, hz, op, so);
        ey = ey_n;
        ez = ez_n;
    }
}

template <typename F>
__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_update_h(F *__restrict__ H, const F *__restrict__ E, PsiDev ps, GridDev g, float ch,
           PlaneSrcDev pw, const int *__restrict__ stp)
{
    const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, g.nblocks));
    if (!tc.ok) return;
    update_h_tile<F>(H, E, ps, g, ch, pw, stp, tc);
}

// ------------------------------------------------------------------------------------------
// E whole step (reference update_e / update_e_lut, `openem/kernels/yee.cu:154-337`; complex
// update_e_c, `openem/kernels/bloch.cu:128-221`) fused with the E side of the plane-wave TF/SF
// correction (`openem/kernels/source_dft.cu:58-73`).  The class LUT is staged in LDS and the
// next plane's class index is prefetched, so the per-cell coefficient gather never waits on a
// global load issued in the same iteration.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool lut_in_lds(const MatDev &mat) {
    return mat.mode == 0 && mat.n_classes <= OEM_LDS_CLASSES;
}

// Stage the class LUT (2 float4 per class) in LDS; the caller brackets it with barriers.
__device__ __forceinline__ void stage_lut(float4 *s_lut, const MatDev &mat) {
    const int tid = threadIdx.y * OEM_TK + threadIdx.x;
    const float4 *src = reinterpret_cast<const float4 *>(mat.lut);
    for (int q = tid; q < 2 * mat.n_classes; q += OEM_TK * OEM_TJ) s_lut[q] = src[q];
}

template <typename F>
__device__ __forceinline__ void update_e_tile(F *__restrict__ E, const F *__restrict__ H,
                                              const PsiDev &ps, const GridDev &g,
                                              const MatDev &mat, float cw, const PlaneSrcDev &pw,
                                              const int *__restrict__ stp, const TileCoord &tc,
                                              const float4 *s_lut, bool lds_lut)
{
    typedef FT<F> T;
    const int j = tc.j, k = tc.k;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const size_t ncell = (size_t)g.plane * nx;
    const uint32_t cbytes = (uint32_t)(ncell * T::B);
    const rsrc_t rEx = mk_rsrc(E, cbytes), rEy = mk_rsrc(E + ncell, cbytes),
                 rEz = mk_rsrc(E + 2 * ncell, cbytes);
    const rsrc_t rHx = mk_rsrc(H, cbytes), rHy = mk_rsrc(H + ncell, cbytes),
                 rHz = mk_rsrc(H + 2 * ncell, cbytes);
    const rsrc_t rP = mk_rsrc(ps.e, ps.total * T::B);
    const rsrc_t rC = mk_rsrc(mat.cls, (uint32_t)(ncell * 2));

    const uint32_t op = (uint32_t)(j * nz + k) * T::B;
    const uint32_t oj = (uint32_t)(Y.prv[j] * nz + k) * T::B;     // (j-1, k)
    const uint32_t ok = (uint32_t)(j * nz + Z.prv[k]) * T::B;     // (j, k-1)
    const float mj = axv(Y, j, AX_MPV), mk = axv(Z, k, AX_MPV);
    const float mji = axv(Y, j, AX_MPVI), mki = axv(Z, k, AX_MPVI);
    const float idy = axv(Y, j, AX_IDLD), idz = axv(Z, k, AX_IDLD);
    const float pecy = axv(Y, j, AX_PEC), pecz = axv(Z, k, AX_PEC);
    const int sy = slab_of(j, ny, Y.lo, Y.hi);
    const int sz = slab_of(k, nz, Z.lo, Z.hi);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    float ay = 0.f, by = 1.f, ky = 1.f, az = 0.f, bz = 1.f, kz = 1.f;
    if (sy >= 0) { ay = axv(Y, j, AX_AE); by = axv(Y, j, AX_BE); ky = axv(Y, j, AX_KE); }
    if (sz >= 0) { az = axv(Z, k, AX_AE); bz = axv(Z, k, AX_BE); kz = axv(Z, k, AX_KE); }
    const uint32_t oy = (uint32_t)(max(sy, 0) * nz + k) * T::B;
    const uint32_t oz = (uint32_t)(j * wz + max(sz, 0)) * T::B;

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
    const rsrc_t rHist = mk_rsrc(mat.hist, 3u * hcomp * T::B);
    const uint32_t oh = hrow ? (uint32_t)((j - mat.hb[2]) * hbz + (k - mat.hb[4])) * T::B : 0u;
    const uint32_t oc = (uint32_t)(j * nz + k) * 2u;               // uint16 class index

    const uint32_t planeB = (uint32_t)g.plane * T::B, plane2 = (uint32_t)g.plane * 2u;
    const uint32_t sp0 = (uint32_t)__builtin_amdgcn_readfirstlane(X.prv[tc.i0]) * planeB;
    F hy_m = T::ld(rHy, op, sp0), hz_m = T::ld(rHz, op, sp0);
    uint32_t cls_next = mat.mode == 0 ? bld_u16(rC, oc, (uint32_t)tc.i0 * plane2) : 0u;
    for (int i = tc.i0; i < tc.i1; ++i) {
        const uint32_t so = (uint32_t)i * planeB;
        const uint32_t cls = cls_next;
        if (mat.mode == 0 && i + 1 < tc.i1) cls_next = bld_u16(rC, oc, (uint32_t)(i + 1) * plane2);
        const F hx = T::ld(rHx, op, so), hy = T::ld(rHy, op, so), hz = T::ld(rHz, op, so);
        const F hx_j = T::ld(rHx, oj, so), hx_k = T::ld(rHx, ok, so);
        const F hy_k = T::ld(rHy, ok, so), hz_j = T::ld(rHz, oj, so);
        F ex = T::ld(rEx, op, so), ey = T::ld(rEy, op, so), ez = T::ld(rEz, op, so);
        const int sx = slab_of(i, nx, X.lo, X.hi);
        const uint32_t sx0 = (ps.off[0] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sx1 = (ps.off[1] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sy0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sy1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sz0 = (ps.off[4] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        const uint32_t sz1 = (ps.off[5] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        F psx0 = T::zero(), psx1 = T::zero(), psy0 = T::zero(), psy1 = T::zero(),
          psz0 = T::zero(), psz1 = T::zero();
        if (sx >= 0) { psx0 = T::ld(rP, op, sx0); psx1 = T::ld(rP, op, sx1); }
        if (sy >= 0) { psy0 = T::ld(rP, oy, sy0); psy1 = T::ld(rP, oy, sy1); }
        if (sz >= 0) { psz0 = T::ld(rP, oz, sz0); psz1 = T::ld(rP, oz, sz1); }

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
        F phx = T::zero(), phy = T::zero(), phz = T::zero();
        if (fl) {
            const uint32_t sh = (uint32_t)(i - mat.hb[0]) * (uint32_t)(hby * hbz) * T::B;
            if (fl & 1u) phx = T::ld(rHist, oh, sh);
            if (fl & 2u) phy = T::ld(rHist, oh, sh + hcomp * T::B);
            if (fl & 4u) phz = T::ld(rHist, oh, sh + 2u * hcomp * T::B);
        }

        const ctab_t xr = ctab_row(X, i);
        const float mi = xr[AX_MPV], mii = xr[AX_MPVI];
        const float idx = xr[AX_IDLD], pecx = xr[AX_PEC];
        const F dzy = (hz - T::mmul(hz_j, mj, mji)) * idy;
        const F dyz = (hy - T::mmul(hy_k, mk, mki)) * idz;
        const F dxz = (hx - T::mmul(hx_k, mk, mki)) * idz;
        const F dzx = (hz - T::mmul(hz_m, mi, mii)) * idx;
        const F dyx = (hy - T::mmul(hy_m, mi, mii)) * idx;
        const F dxy = (hx - T::mmul(hx_j, mj, mji)) * idy;

        F s_ex_y = dzy, s_ez_y = dxy, s_ex_z = dyz, s_ey_z = dxz, s_ey_x = dzx, s_ez_x = dyx;
        if (sx >= 0) {
            const float ax_ = xr[AX_AE], bx_ = xr[AX_BE], kx_ = xr[AX_KE];
            s_ey_x = stretch_v(psx0, dzx, ax_, bx_, kx_);
            s_ez_x = stretch_v(psx1, dyx, ax_, bx_, kx_);
            T::st(rP, psx0, op, sx0);
            T::st(rP, psx1, op, sx1);
        }
        if (sy >= 0) {
            s_ex_y = stretch_v(psy0, dzy, ay, by, ky);
            s_ez_y = stretch_v(psy1, dxy, ay, by, ky);
            T::st(rP, psy0, oy, sy0);
            T::st(rP, psy1, oy, sy1);
        }
        if (sz >= 0) {
            s_ex_z = stretch_v(psz0, dyz, az, bz, kz);
            s_ey_z = stretch_v(psz1, dxz, az, bz, kz);
            T::st(rP, psz0, oz, sz0);
            T::st(rP, psz1, oz, sz1);
        }
        ex = (cax * ex + cbx * (s_ex_y - s_ex_z - cw * phx)) * pecy * pecz;
        ey = (cay * ey + cby * (s_ey_z - s_ey_x - cw * phy)) * pecz * pecx;
        ez = (caz * ez + cbz * (s_ez_x - s_ez_y - cw * phz)) * pecx * pecy;
        if (inj_x) ex = T::addre(ex, cbx * inj_ex);
        if (inj_y) ey = T::addre(ey, cby * inj_ey);
        T::st(rEx, ex, op, so);
        T::st(rEy, ey, op, so);
        T::st(rEz, ez, op, so);
        hy_m = hy;
        hz_m = hz;
    }
}

template <typename F>
__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_update_e(F *__restrict__ E, const F *__restrict__ H, PsiDev ps, GridDev g, MatDev mat,
           float cw, PlaneSrcDev pw, const int *__restrict__ stp)
{
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    const bool lds_lut = lut_in_lds(mat);
    if (lds_lut) {
        stage_lut(s_lut, mat);
        __syncthreads();
    }
    const TileCoord tc = tile_coord(g, xcd_remap(blockIdx.x, g.nblocks));
    if (!tc.ok) return;
    update_e_tile<F>(E, H, ps, g, mat, cw, pw, stp, tc, s_lut, lds_lut);
}

// Plane-wave H-side correction at cell k (the injection term of update_h).
__device__ __forceinline__ void pw_h_inj(const PlaneSrcDev &pw, int k, const int *stp,
                                         float &inj_hx, float &inj_hy) {
    inj_hx = 0.f;
    inj_hy = 0.f;
    if (!pw.n) return;
    const int n = step_now(stp);
    for (int s = 0; s < pw.n; ++s)
        if (k == pw.kh[s]) {
            const float v = pw.coef_h[s] * (float)pw.ex_inc[s][n];
            if (pw.pol[s] == 0) inj_hy += v; else inj_hx += v;
        }
}

// ------------------------------------------------------------------------------------------
// Fused H + E step.  A work-group owns a tile of OEM_CR rows (j) x 64 cells (k)
// and marches it along x; every thread of the OEM_CR row waves keeps one (j, k) column, as in
// the separate sweeps, and computes H^{n+1/2} and then E^{n+1} of each plane:
//   * the old Ey, Ez of the plane and the new Hy, Hz of the plane before are carried in
//     registers (the x neighbours of H and of E);
//   * the new H of the row above (j - 1) and of the lane to the left (k - 1) come from LDS, one
//     barrier per plane, double-buffered by plane parity;
//   * one more wave recomputes, from the same old values with the same arithmetic, the new H of
//     the halo row Y.prv[j0] and of the halo column Z.prv[k0] (not stored): the values are
//     bit-identical to what the owning work-groups store.
// Old and new fields are separate buffers (ping-pong between steps), so no work-group reads a
// value another one writes.  Each field is read once and written once per step.
// ------------------------------------------------------------------------------------------
#ifndef OEM_CR
#define OEM_CR 7            // owned rows (waves) per work-group; one halo wave more
#endif
// s_waitcnt vmcnt(0) with the other counters left alone (gfx9 encoding: vmcnt bits 3:0 and
// 15:14, expcnt 6:4, lgkmcnt 11:8)
#define OEM_WAIT_VM0 0x0F70

// Buffer descriptors of the fused step: a field set is E then H in one allocation, a psi set
// the H side then the E side; one descriptor each keeps the SGPRs the march needs low.
struct ColIO {
    rsrc_t fo, fn;              // old / new field set
    rsrc_t po, pn;              // old / new psi set
    uint32_t cb;                // bytes per field component
    uint32_t hof;               // bytes from E to H in a field set (3 * cb)
    uint32_t pef;               // bytes from the H-side to the E-side psi in a psi set
};

// The H update at one (j, k) column: its per-column constants and the old Ey, Ez of the
// current plane.
template <typename F>
struct HColumn {
    uint32_t op, oj, ok;        // byte offsets in the plane: (j, k), (Y.nxt[j], k), (j, Z.nxt[k])
    uint32_t oy, oz;            // psi offsets in the y / z slab planes
    int sy, sz;                 // slab indices (-1 outside)
    float mj, mji, idy, mk, mki, idz;
    float ay, by, ky, az, bz, kz;
    float ihx, ihy;             // plane-wave H corrections at k
    F ey, ez;
};

template <typename F>
__device__ __forceinline__ void hcolumn_init(HColumn<F> &c, const GridDev &g,
                                             const PlaneSrcDev &pw, const int *stp, int j, int k,
                                             int jn, int kn) {
    typedef FT<F> T;
    const AxisDev &Y = g.ax[1], &Z = g.ax[2];
    const int nz = g.nz, wz = Z.lo + Z.hi;
    c.op = (uint32_t)(j * nz + k) * T::B;
    c.oj = (uint32_t)(jn * nz + k) * T::B;
    c.ok = (uint32_t)(j * nz + kn) * T::B;
    c.sy = slab_of(j, g.ny, Y.lo, Y.hi);
    c.sz = slab_of(k, nz, Z.lo, Z.hi);
    c.oy = (uint32_t)(max(c.sy, 0) * nz + k) * T::B;
    c.oz = (uint32_t)(j * wz + max(c.sz, 0)) * T::B;
    c.mj = axv(Y, j, AX_MNX); c.mji = axv(Y, j, AX_MNXI); c.idy = axv(Y, j, AX_IDLP);
    c.mk = axv(Z, k, AX_MNX); c.mki = axv(Z, k, AX_MNXI); c.idz = axv(Z, k, AX_IDLP);
    c.ay = 0.f; c.by = 1.f; c.ky = 1.f; c.az = 0.f; c.bz = 1.f; c.kz = 1.f;
    if (c.sy >= 0) { c.ay = axv(Y, j, AX_AH); c.by = axv(Y, j, AX_BH); c.ky = axv(Y, j, AX_KH); }
    if (c.sz >= 0) { c.az = axv(Z, k, AX_AH); c.bz = axv(Z, k, AX_BH); c.kz = axv(Z, k, AX_KH); }
    pw_h_inj(pw, k, stp, c.ihx, c.ihy);
    c.ey = c.ez = T::zero();
}

// Old Ey, Ez of plane i into the column.
template <typename F>
__device__ __forceinline__ void hcolumn_load_e(HColumn<F> &c, const ColIO &io, const GridDev &g,
                                               int i) {
    typedef FT<F> T;
    const uint32_t so = (uint32_t)i * ((uint32_t)g.plane * T::B);
    c.ey = T::ld(io.fo, c.op, so + io.cb);
    c.ez = T::ld(io.fo, c.op, so + 2u * io.cb);
}

// The old fields the H update of a column at plane i reads: E at the cell, at (j + 1), (k + 1)
// and X.nxt[i] (Ey, Ez), and H at the cell.
template <typename F>
struct HLoads {
    F ex, ex_j, ex_k, ey_k, ez_j, ey_n, ez_n, h0, h1, h2;
    F psx0, psx1, psy0, psy1, psz0, psz1;      // H-side psi of the cell (zero outside slabs)
};

// Byte offsets of the H-side psi of plane i in a psi set (x slab, y slab, z slab; two each).
struct PsiOffH {
    uint32_t x0, x1, y0, y1, z0, z1;
    int sx;
};
template <typename F>
__device__ __forceinline__ PsiOffH psi_off_h(const GridDev &g, const PsiDev &ps, int i) {
    typedef FT<F> T;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    PsiOffH o;
    o.sx = slab_of(i, g.nx, X.lo, X.hi);
    o.x0 = (ps.off[0] + (uint32_t)max(o.sx, 0) * (uint32_t)g.plane) * T::B;
    o.x1 = (ps.off[1] + (uint32_t)max(o.sx, 0) * (uint32_t)g.plane) * T::B;
    o.y0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * g.nz)) * T::B;
    o.y1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * g.nz)) * T::B;
    o.z0 = (ps.off[4] + (uint32_t)i * (uint32_t)(g.ny * wz)) * T::B;
    o.z1 = (ps.off[5] + (uint32_t)i * (uint32_t)(g.ny * wz)) * T::B;
    return o;
}

template <typename F>
__device__ __forceinline__ void hcolumn_load(HLoads<F> &v, const HColumn<F> &c, const ColIO &io,
                                             const GridDev &g, const PsiDev &ps, int i, int in_) {
    typedef FT<F> T;
    const uint32_t planeB = (uint32_t)g.plane * T::B;
    const uint32_t so = (uint32_t)i * planeB, sn = (uint32_t)in_ * planeB, sh = so + io.hof;
    const uint32_t c1 = io.cb, c2 = 2u * io.cb;
    const PsiOffH po = psi_off_h<F>(g, ps, i);
    v.psx0 = v.psx1 = v.psy0 = v.psy1 = v.psz0 = v.psz1 = T::zero();
    if (po.sx >= 0) { v.psx0 = T::ld(io.po, c.op, po.x0); v.psx1 = T::ld(io.po, c.op, po.x1); }
    if (c.sy >= 0) { v.psy0 = T::ld(io.po, c.oy, po.y0); v.psy1 = T::ld(io.po, c.oy, po.y1); }
    if (c.sz >= 0) { v.psz0 = T::ld(io.po, c.oz, po.z0); v.psz1 = T::ld(io.po, c.oz, po.z1); }
    v.ex = T::ld(io.fo, c.op, so);
    v.ex_j = T::ld(io.fo, c.oj, so);
    v.ex_k = T::ld(io.fo, c.ok, so);
    v.ey_k = T::ld(io.fo, c.ok, so + c1);
    v.ez_j = T::ld(io.fo, c.oj, so + c2);
    v.ey_n = T::ld(io.fo, c.op, sn + c1);
    v.ez_n = T::ld(io.fo, c.op, sn + c2);
    v.h0 = T::ld(io.fo, c.op, sh);
    v.h1 = T::ld(io.fo, c.op, sh + c1);
    v.h2 = T::ld(io.fo, c.op, sh + c2);
}

// H^{n+1/2} of the column at plane i from its loaded inputs (the arithmetic of update_h_tile);
// `xr` is the x-axis table row of plane i.  The column's Ey, Ez move on to plane X.nxt[i].  With
// `store` the new H and H-side psi are written.
template <typename F>
__device__ __forceinline__ void hcolumn_do(HColumn<F> &c, const HLoads<F> &v, const ColIO &io,
                                           const GridDev &g, const PsiDev &ps, int i,
                                           ctab_t xr, float ch, bool store, F &hx, F &hy,
                                           F &hz) {
    typedef FT<F> T;
    const uint32_t so = (uint32_t)i * ((uint32_t)g.plane * T::B), sh = so + io.hof;
    const uint32_t c1 = io.cb, c2 = 2u * io.cb;
    const PsiOffH po = psi_off_h<F>(g, ps, i);
    const int sx = po.sx;
    const uint32_t sx0 = po.x0, sx1 = po.x1, sy0 = po.y0, sy1 = po.y1, sz0 = po.z0, sz1 = po.z1;
    F psx0 = v.psx0, psx1 = v.psx1, psy0 = v.psy0, psy1 = v.psy1, psz0 = v.psz0, psz1 = v.psz1;

    const float mi = xr[AX_MNX], mii = xr[AX_MNXI], idx = xr[AX_IDLP];
    const F dzy = (T::mmul(v.ez_j, c.mj, c.mji) - c.ez) * c.idy;
    const F dyz = (T::mmul(v.ey_k, c.mk, c.mki) - c.ey) * c.idz;
    const F dxz = (T::mmul(v.ex_k, c.mk, c.mki) - v.ex) * c.idz;
    const F dzx = (T::mmul(v.ez_n, mi, mii) - c.ez) * idx;
    const F dyx = (T::mmul(v.ey_n, mi, mii) - c.ey) * idx;
    const F dxy = (T::mmul(v.ex_j, c.mj, c.mji) - v.ex) * c.idy;
    F s_hx_y = dzy, s_hz_y = dxy, s_hx_z = dyz, s_hy_z = dxz, s_hy_x = dzx, s_hz_x = dyx;
    if (sx >= 0) {
        const float a_ = xr[AX_AH], b_ = xr[AX_BH], k_ = xr[AX_KH];
        s_hy_x = stretch_v(psx0, dzx, a_, b_, k_);
        s_hz_x = stretch_v(psx1, dyx, a_, b_, k_);
        if (store) { T::st(io.pn, psx0, c.op, sx0); T::st(io.pn, psx1, c.op, sx1); }
    }
    if (c.sy >= 0) {
        s_hx_y = stretch_v(psy0, dzy, c.ay, c.by, c.ky);
        s_hz_y = stretch_v(psy1, dxy, c.ay, c.by, c.ky);
        if (store) { T::st(io.pn, psy0, c.oy, sy0); T::st(io.pn, psy1, c.oy, sy1); }
    }
    if (c.sz >= 0) {
        s_hx_z = stretch_v(psz0, dyz, c.az, c.bz, c.kz);
        s_hy_z = stretch_v(psz1, dxz, c.az, c.bz, c.kz);
        if (store) { T::st(io.pn, psz0, c.oz, sz0); T::st(io.pn, psz1, c.oz, sz1); }
    }
    hx = T::addre(v.h0 - ch * (s_hx_y - s_hx_z), c.ihx);
    hy = T::addre(v.h1 - ch * (s_hy_z - s_hy_x), c.ihy);
    hz = v.h2 - ch * (s_hz_x - s_hz_y);
    if (store) {
        T::st(io.fn, hx, c.op, sh);
        T::st(io.fn, hy, c.op, sh + c1);
        T::st(io.fn, hz, c.op, sh + c2);
    }
    c.ey = v.ey_n;
    c.ez = v.ez_n;
}

// DFT monitors whose snapshot the fused step writes as it stores the new fields, in the layout
// of k_dft_snap (a flux plane: E_t1, E_t2 and the H averages over the plane and the one below;
// a field box: E and H); no work-group meets more than three of them.  Cells that an E-side point list
// corrects after the step are snapped again afterwards (k_dft_snap_list).
#define OEM_FUSE_MONS 4
struct FuseMon {
    int kind, axis, km;         // kind 0: a plane (normal axis, index); 1: a box
    int o[3], b[3];             // box
    int npts, T;
    void *snap;
};
struct FuseSnap {
    int n;
    FuseMon m[OEM_FUSE_MONS];
};

// The monitors of `fs` (bit t: monitor t) whose cells meet the box [i0, i1) x [j0, j1) x [k0, k1).
__device__ __forceinline__ uint32_t fuse_snap_mask(const FuseSnap &fs, int nx, int ny, int nz,
                                                   int i0, int i1, int j0, int j1, int k0, int k1)
{
    uint32_t mask = 0;
    for (int t = 0; t < fs.n; ++t) {
        const FuseMon m = fs.m[t];
        int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        if (m.kind == 1) {
            for (int a = 0; a < 3; ++a) { lo[a] = m.o[a]; hi[a] = m.o[a] + m.b[a]; }
        } else {
            lo[m.axis] = m.km;
            hi[m.axis] = m.km + 1;
        }
        if (lo[0] < i1 && hi[0] > i0 && lo[1] < j1 && hi[1] > j0 && lo[2] < k1 && hi[2] > k0)
            mask |= 1u << t;
    }
    return mask;
}

// One monitor's snapshot for the columns of a work-group, set up before the march: point
// q = (i - ilo) * si + qjk of the plane i in [ilo, ihi) (qjk < 0: the column is not on the
// monitor); `sel` 0: a field box, 1 + a: a flux plane normal to axis a.
template <typename F>
struct SnapPlan {
    F *row;          // the ring row of this step
    int np, sel, ilo, ihi, si, qjk;
};

template <typename F>
__device__ __forceinline__ SnapPlan<F> snap_plan(const FuseMon &m, int step, int nx, int ny,
                                                 int nz, int j, int k)
{
    SnapPlan<F> p;
    p.np = m.npts;
    if (m.kind == 1) {
        p.sel = 0;
        p.ilo = m.o[0]; p.ihi = m.o[0] + m.b[0]; p.si = m.b[1] * m.b[2];
        p.qjk = ((unsigned)(j - m.o[1]) < (unsigned)m.b[1] && (unsigned)(k - m.o[2]) < (unsigned)m.b[2])
                    ? (j - m.o[1]) * m.b[2] + (k - m.o[2]) : -1;
    } else {
        p.sel = 1 + m.axis;
        p.ilo = 0; p.ihi = nx;
        if (m.axis == 0) { p.ilo = m.km; p.ihi = m.km + 1; p.si = 0; p.qjk = j * nz + k; }
        else if (m.axis == 1) { p.si = nz; p.qjk = j == m.km ? k : -1; }
        else { p.si = ny; p.qjk = k == m.km ? j : -1; }
    }
    p.row = static_cast<F *>(m.snap) + (size_t)(step % m.T) * (m.kind == 1 ? 6 : 4) * (size_t)m.npts;
    return p;
}

// The snapshot values of one cell under plan p, to the ring (stage == nullptr) or to slot `slot`
// of an LDS stage laid out [component][64].
template <typename F>
__device__ __forceinline__ void snap_planned(const SnapPlan<F> &p, int i, F *stage, int slot, F ex,
                                             F ey, F ez, F hx, F hy, F hz, F hym_i, F hzm_i,
                                             F hxm_j, F hzm_j, F hxm_k, F hym_k)
{
    size_t np = 64, q = (size_t)slot;
    F *row = stage;
    if (!stage) {
        np = (size_t)p.np;
        q = (size_t)((i - p.ilo) * p.si + p.qjk);
        row = p.row;
    }
    if (p.sel == 0) {
        row[q] = ex; row[np + q] = ey; row[2 * np + q] = ez;
        row[3 * np + q] = hx; row[4 * np + q] = hy; row[5 * np + q] = hz;
    } else if (p.sel == 1) {
        row[q] = ey; row[np + q] = ez;
        row[2 * np + q] = 0.5f * (hy + hym_i); row[3 * np + q] = 0.5f * (hz + hzm_i);
    } else if (p.sel == 2) {
        row[q] = ex; row[np + q] = ez;
        row[2 * np + q] = 0.5f * (hx + hxm_j); row[3 * np + q] = 0.5f * (hz + hzm_j);
    } else {
        row[q] = ex; row[np + q] = ey;
        row[2 * np + q] = 0.5f * (hx + hxm_k); row[3 * np + q] = 0.5f * (hy + hym_k);
    }
}

// Rank the row waves' snapshot points (`part`) across the work-group, one barrier that every
// wave takes: the number of points if at most 64 (each point's slot and ring offset qjk go to
// `slot` and fq[slot]), else -1 (the row waves store their points themselves).
__device__ __forceinline__ int snap_rank(bool part, int w, int lane, int qjk, int *s_fcnt,
                                         int *s_fq, int &slot)
{
    const uint64_t bal = __ballot(part);
    if (w < OEM_CR && lane == 0) s_fcnt[w] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int r = 0; r < OEM_CR; ++r) {
        const int c = s_fcnt[r];
        off += r < w ? c : 0;
        tot += c;
    }
    if (tot > 64) return -1;
    if (part) {
        slot = off + __popcll(bal & ((1ull << lane) - 1ull));
        s_fq[slot] = qjk;
    }
    return __builtin_amdgcn_readfirstlane(tot);
}

// The staged snapshot of plane i (cnt points, ring offsets fq) to the ring, by one wave.
template <typename F>
__device__ __forceinline__ void snap_flush(const SnapPlan<F> &p, const F *stage,
                                           const int *fq, int cnt, int i, int lane)
{
    if (lane >= cnt) return;
    const int nc = p.sel == 0 ? 6 : 4;
    const size_t np = (size_t)p.np, q = (size_t)((i - p.ilo) * p.si + fq[lane]);
    for (int c = 0; c < nc; ++c) p.row[c * np + q] = stage[c * 64 + lane];
}

// MK: how the march reads the material coefficients (a kernel of its own each, so the wait
// analysis of one form is not merged with the loads of another): 0 the class LUT staged in
// LDS, 1 the class LUT in global memory (more classes than LDS holds), 2 dense per-cell arrays.
// MULTI: up to three fused monitors meet a work-group (else at most one, and the kernel holds
// no registers for the others).
template <typename F, int MK, bool MULTI>
__global__ void __launch_bounds__(64 * (OEM_CR + 1))
k_update_he(const F *__restrict__ Eo, const F *__restrict__ Ho, F *__restrict__ En,
                F *__restrict__ Hn, PsiDev ps, PsiDev pn, GridDev g, MatDev mat, float ch,
                float cw, PlaneSrcDev pw, const int *__restrict__ stp, FuseSnap fs)
{
    typedef FT<F> T;
    // Ho = Eo + 3 components and ps.e = ps.h + ps.total (the host allocates them so)
    ColIO io;
    {
        const size_t nc = (size_t)g.plane * g.nx;
        io.cb = (uint32_t)(nc * T::B);
        io.hof = 3u * io.cb;
        io.pef = ps.total * T::B;
        io.fo = mk_rsrc(Eo, 2u * io.hof); io.fn = mk_rsrc(En, 2u * io.hof);
        io.po = mk_rsrc(ps.h, 2u * io.pef); io.pn = mk_rsrc(pn.h, 2u * io.pef);
    }
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    // new H per row slot (0: the halo row, 1 + w: row wave w) and of the halo column, by
    // plane parity
    __shared__ F s_row[2][3][OEM_CR + 1][64];
    __shared__ F s_col[2][2][OEM_CR];
    // snapshot stage by plane parity ([component][point]), the points' ring offsets, and the
    // row waves' point counts
    __shared__ F s_fsn[2][6 * 64];
    __shared__ int s_fq[64];
    __shared__ int s_fcnt[OEM_CR];
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int nkt = (nz + 63) / 64;
    // per-work-group quantities through readfirstlane: the compiler then keeps them in SGPRs
    const int lb = xcd_remap(blockIdx.x, g.nblocks);
    const int tile = __builtin_amdgcn_readfirstlane(lb % g.tiles);
    const int chunk = __builtin_amdgcn_readfirstlane(lb / g.tiles);
    const int j0 = __builtin_amdgcn_readfirstlane((tile / nkt) * OEM_CR);
    const int k0 = __builtin_amdgcn_readfirstlane((tile % nkt) * 64);
    const int i0 = chunk * g.chunk, i1 = min(i0 + g.chunk, nx);
    constexpr bool lds_lut = MK == 0;
    if constexpr (lds_lut) {
        const int tid = threadIdx.y * 64 + threadIdx.x;
        const float4 *src = reinterpret_cast<const float4 *>(mat.lut);
        for (int q = tid; q < 2 * mat.n_classes; q += 64 * (OEM_CR + 1)) s_lut[q] = src[q];
        __syncthreads();
    }
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lane = threadIdx.x;
    const int xn_last = __builtin_amdgcn_readfirstlane(X.nxt[nx - 1]);
    const int yn_last = __builtin_amdgcn_readfirstlane(Y.nxt[ny - 1]);
    const int zn_last = __builtin_amdgcn_readfirstlane(Z.nxt[nz - 1]);
    const uint32_t planeB = (uint32_t)g.plane * T::B, plane2 = (uint32_t)g.plane * 2u;
    const size_t ncell = (size_t)g.plane * nx;

    // Threads outside the grid work on a clamped in-range column and store nothing.
    const int kl = min(k0 + lane, nz - 1);
    const int xp0 = __builtin_amdgcn_readfirstlane(X.prv[i0]);
    // the fused DFT monitor that meets this work-group's tile and chunk (the host fuses monitors
    // whose work-groups are disjoint; most work-groups meet none), set up here so that the march
    // reads no descriptor.  Up to 64 points go through an LDS stage that the halo wave stores
    // with full-width stores after the next barrier (a row wave holds only a few of them: 2
    // lanes of a 2-layer box, 1 of a z plane)
    // A second and third monitor meeting the work-group (the host lets at most three meet any
    // one) have their points stored directly by the row waves (fp2, fp3).
    int fcnt = -1, fslot = -1;
    bool fplan = false;
    SnapPlan<F> fp{}, fp2{}, fp3{};
    fp.ilo = fp.ihi = 0;
    fp2.ilo = fp2.ihi = 0;
    fp3.ilo = fp3.ihi = 0;
    if (fs.n) {
        uint32_t fmask = __builtin_amdgcn_readfirstlane(
            fuse_snap_mask(fs, nx, ny, nz, i0, i1, j0, j0 + OEM_CR, k0, k0 + 64));
        if (fmask) {
            const int fstep = __builtin_amdgcn_readfirstlane(step_now(stp));
            const int jr = min(j0 + min(w, OEM_CR - 1), ny - 1);
            // all but qjk are the same across the work-group
            auto uniform = [](SnapPlan<F> &p) {
                const uint64_t rp = reinterpret_cast<uint64_t>(p.row);
                const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)rp);
                const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(rp >> 32));
                p.row = reinterpret_cast<F *>(((uint64_t)hi << 32) | lo);
                p.np = __builtin_amdgcn_readfirstlane(p.np);
                p.sel = __builtin_amdgcn_readfirstlane(p.sel);
                p.ilo = __builtin_amdgcn_readfirstlane(p.ilo);
                p.ihi = __builtin_amdgcn_readfirstlane(p.ihi);
                p.si = __builtin_amdgcn_readfirstlane(p.si);
            };
            fp = snap_plan<F>(fs.m[__builtin_ctz(fmask)], fstep, nx, ny, nz, jr, kl);
            uniform(fp);
            fplan = true;
            fmask &= fmask - 1u;
            if (MULTI && fmask) {
                fp2 = snap_plan<F>(fs.m[__builtin_ctz(fmask)], fstep, nx, ny, nz, jr, kl);
                uniform(fp2);
                fmask &= fmask - 1u;
                if (fmask) {
                    fp3 = snap_plan<F>(fs.m[__builtin_ctz(fmask)], fstep, nx, ny, nz, jr, kl);
                    uniform(fp3);
                }
            }
        }
    }
    if (w == OEM_CR) {
        // the halo wave: the new H of (Y.prv[j0], k0 + lane) and of (j0 + lane, Z.prv[k0]) for
        // lane < OEM_CR, into LDS slot 0 and the column slots; one barrier per plane, as the
        // row waves
        const int jh = __builtin_amdgcn_readfirstlane(Y.prv[j0]);
        const int kh = __builtin_amdgcn_readfirstlane(Z.prv[k0]);
        const int jc = min(j0 + min(lane, OEM_CR - 1), ny - 1);
        HColumn<F> cr, cc;
        hcolumn_init<F>(cr, g, pw, stp, jh, kl, next_of(jh, ny, yn_last), next_of(kl, nz, zn_last));
        hcolumn_init<F>(cc, g, pw, stp, jc, kh, next_of(jc, ny, yn_last), next_of(kh, nz, zn_last));
        hcolumn_load_e<F>(cr, io, g, i0);
        hcolumn_load_e<F>(cc, io, g, i0);
        if (fplan) fcnt = snap_rank(false, w, lane, 0, s_fcnt, s_fq, fslot);
#pragma unroll 1
        for (int i = i0; i < i1; ++i) {
            const int buf = (i - i0) & 1;
            const int in_ = next_of(i, nx, xn_last);
            F hx, hy, hz;
            HLoads<F> vr, vc;
            hcolumn_load<F>(vr, cr, io, g, ps, i, in_);
            hcolumn_load<F>(vc, cc, io, g, ps, i, in_);
            const ctab_t xr = ctab_row(X, i);
            hcolumn_do<F>(cr, vr, io, g, ps, i, xr, ch, false, hx, hy, hz);
            s_row[buf][0][0][lane] = hx;
            s_row[buf][2][0][lane] = hz;
            hcolumn_do<F>(cc, vc, io, g, ps, i, xr, ch, false, hx, hy, hz);
            if (lane < OEM_CR) {
                s_col[buf][0][lane] = hx;
                s_col[buf][1][lane] = hy;
            }
            __syncthreads();
            // the row waves staged plane i - 1 before this barrier
            if (fcnt > 0 && i > i0 && i - 1 >= fp.ilo && i - 1 < fp.ihi)
                snap_flush<F>(fp, s_fsn[buf ^ 1], s_fq, fcnt, i - 1, lane);
        }
        if (fcnt > 0) {
            __syncthreads();
            if (i1 - 1 >= fp.ilo && i1 - 1 < fp.ihi)
                snap_flush<F>(fp, s_fsn[(i1 - 1 - i0) & 1], s_fq, fcnt, i1 - 1, lane);
        }
        return;
    }

    // a row wave: column (j0 + w, k0 + lane)
    const int j = min(j0 + w, ny - 1), k = kl;
    const bool own = j0 + w < ny && k0 + lane < nz;
    HColumn<F> ca;
    hcolumn_init<F>(ca, g, pw, stp, j, k, next_of(j, ny, yn_last), next_of(k, nz, zn_last));
    const float mjE = axv(Y, j, AX_MPV), mjEi = axv(Y, j, AX_MPVI);
    const float idyE = axv(Y, j, AX_IDLD), pecy = axv(Y, j, AX_PEC);
    const float mkE = axv(Z, k, AX_MPV), mkEi = axv(Z, k, AX_MPVI);
    const float idzE = axv(Z, k, AX_IDLD), pecz = axv(Z, k, AX_PEC);
    float ayE = 0.f, byE = 1.f, kyE = 1.f, azE = 0.f, bzE = 1.f, kzE = 1.f;
    if (ca.sy >= 0) { ayE = axv(Y, j, AX_AE); byE = axv(Y, j, AX_BE); kyE = axv(Y, j, AX_KE); }
    if (ca.sz >= 0) { azE = axv(Z, k, AX_AE); bzE = axv(Z, k, AX_BE); kzE = axv(Z, k, AX_KE); }
    float inj_ex = 0.f, inj_ey = 0.f;
    bool inj_x = false, inj_y = false;
    if (pw.n) {
        const int n = step_now(stp);
        for (int s = 0; s < pw.n; ++s)
            if (k == pw.ks[s]) {
                const float v = pw.coef_e[s] * (float)pw.hy_inc[s][n];
                if (pw.pol[s] == 0) { inj_ex += v; inj_x = true; } else { inj_ey += v; inj_y = true; }
            }
    }
    const rsrc_t rC = mk_rsrc(mat.cls, (uint32_t)(ncell * 2));
    const int hbx = mat.hb[1] - mat.hb[0], hby = mat.hb[3] - mat.hb[2], hbz = mat.hb[5] - mat.hb[4];
    const uint32_t hcomp = (uint32_t)(hbx * hby * hbz);
    const rsrc_t rHist = mk_rsrc(mat.hist, 3u * hcomp * T::B);
    const bool hrow = mat.hist && own && j >= mat.hb[2] && j < mat.hb[3] && k >= mat.hb[4] &&
                      k < mat.hb[5];
    const uint32_t oh = hrow ? (uint32_t)((j - mat.hb[2]) * hbz + (k - mat.hb[4])) * T::B : 0u;
    const uint32_t oc = (uint32_t)(j * nz + k) * 2u;

    // the new Hy, Hz of the plane before i0 (the x neighbours of E at i0), not stored
    F hy_m, hz_m;
    {
        HColumn<F> cp = ca;
        hcolumn_load_e<F>(cp, io, g, xp0);
        HLoads<F> vp;
        hcolumn_load<F>(vp, cp, io, g, ps, xp0, next_of(xp0, nx, xn_last));
        F hx_;
        hcolumn_do<F>(cp, vp, io, g, ps, xp0, ctab_row(X, xp0), ch, false, hx_, hy_m,
                      hz_m);
    }
    hcolumn_load_e<F>(ca, io, g, i0);
    uint32_t cls_c = MK != 2 ? bld_u16(rC, oc, (uint32_t)i0 * plane2) : 0u;
    if (fplan) fcnt = snap_rank(own && fp.qjk >= 0, w, lane, fp.qjk, s_fcnt, s_fq, fslot);

#pragma unroll 1
    for (int i = i0; i < i1; ++i) {
        const int buf = (i - i0) & 1;
        const int in_ = next_of(i, nx, xn_last);
        const uint32_t so = (uint32_t)i * planeB;
        const int sx = slab_of(i, nx, X.lo, X.hi);
        const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
        const uint32_t sx0 = (ps.off[0] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sx1 = (ps.off[1] + (uint32_t)max(sx, 0) * (uint32_t)g.plane) * T::B;
        const uint32_t sy0 = (ps.off[2] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sy1 = (ps.off[3] + (uint32_t)i * (uint32_t)(wy * nz)) * T::B;
        const uint32_t sz0 = (ps.off[4] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        const uint32_t sz1 = (ps.off[5] + (uint32_t)i * (uint32_t)(ny * wz)) * T::B;
        const bool hin = hrow && i >= mat.hb[0] && i < mat.hb[1];
        // the update coefficients of the cell's class (its index loaded a plane ahead); the ADE
        // history is read only where the class flags a dispersive component (the box around
        // the dispersive cells is mostly other material)
        const uint32_t cls = cls_c;
        float cax, cbx, cay, cby, caz, cbz;
        uint32_t fl;
        if constexpr (MK == 0) {
            const float4 r0 = s_lut[2 * cls], r1 = s_lut[2 * cls + 1];
            cax = r0.x; cbx = r0.y; cay = r0.z; cby = r0.w; caz = r1.x; cbz = r1.y;
            fl = __float_as_uint(r1.z);
        } else if constexpr (MK == 1) {
            const float4 *rw = reinterpret_cast<const float4 *>(mat.lut + 8u * cls);
            const float4 r0 = rw[0], r1 = rw[1];
            cax = r0.x; cbx = r0.y; cay = r0.z; cby = r0.w; caz = r1.x; cbz = r1.y;
            fl = __float_as_uint(r1.z);
        } else {
            const size_t id = (size_t)i * g.plane + (size_t)j * nz + k;
            cax = mat.ca[id]; cay = mat.ca[ncell + id]; caz = mat.ca[2 * ncell + id];
            cbx = mat.cb[id]; cby = mat.cb[ncell + id]; cbz = mat.cb[2 * ncell + id];
            fl = 7u;
        }
        F phx = T::zero(), phy = T::zero(), phz = T::zero();
        if (hin) {
            const uint32_t sh = (uint32_t)(i - mat.hb[0]) * (uint32_t)(hby * hbz) * T::B;
            if (fl & 1u) phx = T::ld(rHist, oh, sh);
            if (fl & 2u) phy = T::ld(rHist, oh, sh + hcomp * T::B);
            if (fl & 4u) phz = T::ld(rHist, oh, sh + 2u * hcomp * T::B);
        }
        F pe0 = T::zero(), pe1 = T::zero(), pe2 = T::zero(), pe3 = T::zero(), pe4 = T::zero(),
          pe5 = T::zero();
        if (sx >= 0) { pe0 = T::ld(io.po, ca.op, sx0 + io.pef); pe1 = T::ld(io.po, ca.op, sx1 + io.pef); }
        if (ca.sy >= 0) { pe2 = T::ld(io.po, ca.oy, sy0 + io.pef); pe3 = T::ld(io.po, ca.oy, sy1 + io.pef); }
        if (ca.sz >= 0) { pe4 = T::ld(io.po, ca.oz, sz0 + io.pef); pe5 = T::ld(io.po, ca.oz, sz1 + io.pef); }
        HLoads<F> cv;
        hcolumn_load<F>(cv, ca, io, g, ps, i, in_);
        if (MK != 2 && i + 1 < i1) cls_c = bld_u16(rC, oc, (uint32_t)(i + 1) * plane2);
        // every input of the plane has arrived before its first store: the E half after the
        // barrier then waits on no vector load (loads and stores share one in-order counter,
        // so such a wait would also wait for this plane's H stores)
        __builtin_amdgcn_s_waitcnt(OEM_WAIT_VM0);
        const F ey0 = ca.ey, ez0 = ca.ez;
        F hx, hy, hz;
        const ctab_t xr = ctab_row(X, i);
        const F ex0 = cv.ex;
        hcolumn_do<F>(ca, cv, io, g, ps, i, xr, ch, own, hx, hy, hz);
        s_row[buf][0][w + 1][lane] = hx;
        s_row[buf][1][w + 1][lane] = hy;
        s_row[buf][2][w + 1][lane] = hz;
        __syncthreads();
        const F hx_j = s_row[buf][0][w][lane], hz_j = s_row[buf][2][w][lane];
        const F hx_k = lane ? s_row[buf][0][w + 1][lane - 1] : s_col[buf][0][w];
        const F hy_k = lane ? s_row[buf][1][w + 1][lane - 1] : s_col[buf][1][w];
        const F hpx = phx, hpy = phy, hpz = phz;
        const float mi = xr[AX_MPV], mii = xr[AX_MPVI];
        const float idx = xr[AX_IDLD], pecx = xr[AX_PEC];
        const F dzy = (hz - T::mmul(hz_j, mjE, mjEi)) * idyE;
        const F dyz = (hy - T::mmul(hy_k, mkE, mkEi)) * idzE;
        const F dxz = (hx - T::mmul(hx_k, mkE, mkEi)) * idzE;
        const F dzx = (hz - T::mmul(hz_m, mi, mii)) * idx;
        const F dyx = (hy - T::mmul(hy_m, mi, mii)) * idx;
        const F dxy = (hx - T::mmul(hx_j, mjE, mjEi)) * idyE;
        F s_ex_y = dzy, s_ez_y = dxy, s_ex_z = dyz, s_ey_z = dxz, s_ey_x = dzx, s_ez_x = dyx;
        if (sx >= 0) {
            const float a_ = xr[AX_AE], b_ = xr[AX_BE], k_ = xr[AX_KE];
            s_ey_x = stretch_v(pe0, dzx, a_, b_, k_);
            s_ez_x = stretch_v(pe1, dyx, a_, b_, k_);
            if (own) { T::st(io.pn, pe0, ca.op, sx0 + io.pef); T::st(io.pn, pe1, ca.op, sx1 + io.pef); }
        }
        if (ca.sy >= 0) {
            s_ex_y = stretch_v(pe2, dzy, ayE, byE, kyE);
            s_ez_y = stretch_v(pe3, dxy, ayE, byE, kyE);
            if (own) { T::st(io.pn, pe2, ca.oy, sy0 + io.pef); T::st(io.pn, pe3, ca.oy, sy1 + io.pef); }
        }
        if (ca.sz >= 0) {
            s_ex_z = stretch_v(pe4, dyz, azE, bzE, kzE);
            s_ey_z = stretch_v(pe5, dxz, azE, bzE, kzE);
            if (own) { T::st(io.pn, pe4, ca.oz, sz0 + io.pef); T::st(io.pn, pe5, ca.oz, sz1 + io.pef); }
        }
        F nex = (cax * ex0 + cbx * (s_ex_y - s_ex_z - cw * hpx)) * pecy * pecz;
        F ney = (cay * ey0 + cby * (s_ey_z - s_ey_x - cw * hpy)) * pecz * pecx;
        F nez = (caz * ez0 + cbz * (s_ez_x - s_ez_y - cw * hpz)) * pecx * pecy;
        if (inj_x) nex = T::addre(nex, cbx * inj_ex);
        if (inj_y) ney = T::addre(ney, cby * inj_ey);
        if (own) {
            T::st(io.fn, nex, ca.op, so);
            T::st(io.fn, ney, ca.op, so + io.cb);
            T::st(io.fn, nez, ca.op, so + 2u * io.cb);
            if (i >= fp.ilo && i < fp.ihi && fp.qjk >= 0)
                snap_planned<F>(fp, i, fcnt >= 0 ? s_fsn[buf] : nullptr, fslot, nex, ney, nez, hx,
                                hy, hz, hy_m, hz_m, hx_j, hz_j, hx_k, hy_k);
            if constexpr (MULTI) {
                if (i >= fp2.ilo && i < fp2.ihi && fp2.qjk >= 0)
                    snap_planned<F>(fp2, i, nullptr, 0, nex, ney, nez, hx, hy, hz, hy_m, hz_m,
                                    hx_j, hz_j, hx_k, hy_k);
                if (i >= fp3.ilo && i < fp3.ihi && fp3.qjk >= 0)
                    snap_planned<F>(fp3, i, nullptr, 0, nex, ney, nez, hx, hy, hz, hy_m, hz_m,
                                    hx_j, hz_j, hx_k, hy_k);
            }
        }
        hy_m = hy;
        hz_m = hz;
    }
    // the halo wave stores the last staged plane after this barrier
    if (fcnt > 0) __syncthreads();
}

// E at the design cells of an adjoint after the step: row (step + 1 - base) of rec (the step read
// from the device counter, so a captured step graph records every step it replays).
template <typename F>
__global__ void __launch_bounds__(256)
k_rec_gather(const F *__restrict__ E, const int *__restrict__ comp,
             const int *__restrict__ cell, int n, size_t ncell, F *__restrict__ rec,
             const int *__restrict__ stp, int64_t base)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t row = (int64_t)step_now(stp) + 1 - base;
    rec[(size_t)row * n + q] = E[(size_t)comp[q] * ncell + cell[q]];
}

// ------------------------------------------------------------------------------------------
// Grouped point corrections: one thread per distinct (component, cell), its points applied in
// input order (no atomics, deterministic).  Used for
//   * point dipoles (reference inject_dipole / inject_dipole_h,
//     `openem/kernels/source_dft.cu:147-206`, which use atomicAdd):
//     F -= scale * coef * amp[src][n]
//   * TF/SF box face corrections (reference tfsf_corr_e / tfsf_corr_h,
//     `openem/kernels/tfsf.cu:49-83`):  F += scale * coef * (table[n][col] * phase)
// scale is ch on the H side and the cell's cb on the E side.
// ------------------------------------------------------------------------------------------
struct PointsDev {
    int n_groups, nsteps1, ncol;
    int e_side;                             // 1: scale by the cell's cb; 0: by ch
    int tfsf;                               // 0: dipole amplitudes, 1: TF/SF table lookups
    int cplx_amp;                           // dipole amplitude entries are complex doubles
    const int *g_comp, *g_cell, *g_ofs;     // per group; ofs (n_groups+1) into the points
    const float *coef;                      // per point
    const int *col;                         // per point: source (dipole) or table column
    const float *phase;                     // per point (re, im) for TF/SF, or nullptr
    const double *tab;                      // dipole: n_src*nsteps1 (x2 if complex);
                                            // TF/SF: nsteps1*ncol complex (re, im)
};

// Group t of a point list at step n.
template <typename F>
__device__ __forceinline__ void point_group(F *__restrict__ Fld, const PointsDev &d,
                                            const MatDev &mat, float ch, size_t ncell, int n,
                                            int t)
{
    typedef FT<F> T;
    if (n >= d.nsteps1) return;
    const int c = d.g_comp[t], id = d.g_cell[t];
    F *Fp = Fld + (size_t)c * ncell;
    float scale;
    if (!d.e_side) {
        scale = ch;
    } else if (mat.mode == 0) {
        scale = mat.lut[8 * (size_t)mat.cls[id] + 2 * c + 1];
    } else {
        scale = mat.cb[(size_t)c * ncell + id];
    }
    F v = Fp[id];
    for (int q = d.g_ofs[t]; q < d.g_ofs[t + 1]; ++q) {
        float ar, ai;
        if (d.tfsf) {
            const double *e = d.tab + 2 * ((size_t)n * d.ncol + d.col[q]);
            const double pr = d.phase[2 * q], pi = d.phase[2 * q + 1];
            ar = (float)(e[0] * pr - e[1] * pi);
            ai = (float)(e[0] * pi + e[1] * pr);
        } else if (d.cplx_amp) {
            const double *e = d.tab + 2 * ((size_t)d.col[q] * d.nsteps1 + n);
            ar = -(float)e[0];
            ai = -(float)e[1];
        } else {
            ar = -(float)d.tab[(size_t)d.col[q] * d.nsteps1 + n];
            ai = 0.f;
        }
        v = T::addc(v, scale * d.coef[q], ar, ai);
    }
    Fp[id] = v;
}

template <typename F>
__global__ void __launch_bounds__(256)
k_points(F *__restrict__ Fld, PointsDev d, MatDev mat, float ch, size_t ncell,
         const int *__restrict__ stp)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.n_groups) return;
    point_group<F>(Fld, d, mat, ch, ncell, step_now(stp), t);
}

// ------------------------------------------------------------------------------------------
// Complex-pole ADE (reference dispersion_pre / dispersion_post / dispersion_step,
// `openem/kernels/dispersion.cu:68-207`): one thread per entry.  On the complex (Bloch) path
// Re E and Im E each drive their own pole state (p / q), since the poles act cell-wise and every
// coefficient of the update is real.
// ------------------------------------------------------------------------------------------
struct DispDev {
    int n_entry;
    const int *comp, *cell, *ofs, *cidx;
    const float *am1_re, *am1_im, *b_re, *b_im;
    float *p_re, *p_im;     // pole state driven by Re E
    float *q_re, *q_im;     // pole state driven by Im E (complex fields only)
    void *hist;             // F elements, (3, bx, by, bz)
    int hb[6];
    // Compact form, used when every entry has the same pole count np (1..4): per entry
    // {comp << 30 | cell, history index}; per pole a 16-bit coefficient index and the state as
    // (re, im) pairs, entry-major (pole q of entry e at e * np + q); coefficients as
    // (A-1 re, A-1 im, B re, B im).  Every input of an entry is then one load away from its
    // index (no CSR offsets, no index division), and the arithmetic is pole_update's.
    int np;                 // 0: the CSR arrays above
    const uint2 *meta;
    const uint16_t *c16;
    const float4 *coef;
    float2 *sp, *sq;        // pole state driven by Re E / by Im E (complex fields only)
};

__device__ __forceinline__ size_t hist_index(const DispDev &d, int c, int cell, int ny, int nz) {
    const int i = cell / (ny * nz), r = cell - i * ny * nz, j = r / nz, k = r - j * nz;
    const int bx = d.hb[1] - d.hb[0], by = d.hb[3] - d.hb[2], bz = d.hb[5] - d.hb[4];
    return (((size_t)c * bx + (i - d.hb[0])) * by + (j - d.hb[2])) * bz + (k - d.hb[4]);
}

// one pole: mode 0 = pre (P += dP with dP = (A-1)P + B e, returns 2 Re dP), 1 = post
// (P += B e), 2 = post then pre with the same e (the fused step).
__device__ __forceinline__ float pole_update(float *pr_, float *pi_, int t, int ci,
                                            const DispDev &d, float e, int mode) {
    const float br = d.b_re[ci], bi = d.b_im[ci];
    float pr = pr_[t], pi = pi_[t];
    if (mode >= 1) {
        pr = pr + br * e;
        pi = pi + bi * e;
        if (mode == 1) {
            pr_[t] = pr;
            pi_[t] = pi;
            return 0.f;
        }
    }
    const float ar = d.am1_re[ci], ai = d.am1_im[ci];
    const float dr = ar * pr - ai * pi + br * e;
    const float di = ar * pi + ai * pr + bi * e;
    pr_[t] = pr + dr;
    pi_[t] = pi + di;
    return 2.0f * dr;
}

__device__ __forceinline__ void hist_store(void *h, size_t q, float w, float) {
    static_cast<float *>(h)[q] = w;
}

// pole_update on a (re, im) state pair with its coefficients (A-1 re, A-1 im, B re, B im).
__device__ __forceinline__ float pole_pair(float2 &s, const float4 cf, float e, int mode) {
    const float br = cf.z, bi = cf.w;
    float pr = s.x, pi = s.y;
    if (mode >= 1) {
        pr = pr + br * e;
        pi = pi + bi * e;
        if (mode == 1) {
            s = make_float2(pr, pi);
            return 0.f;
        }
    }
    const float ar = cf.x, ai = cf.y;
    const float dr = ar * pr - ai * pi + br * e;
    const float di = ar * pi + ai * pr + bi * e;
    s = make_float2(pr + dr, pi + di);
    return 2.0f * dr;
}

// Entry e of the compact form with NP poles: every load first, then the updates in pole order
// (the sums round as the CSR form's loop does).
template <typename F, int NP>
__device__ __forceinline__ void disp_entry_np(const DispDev &d, const F *__restrict__ E,
                                              size_t ncell, int mode, int e)
{
    const uint2 mt = d.meta[e];
    const uint32_t c = mt.x >> 30, id = mt.x & 0x3fffffffu;
    const size_t t0 = (size_t)e * NP;
    float2 sp[NP];
    float4 cf[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        sp[q] = d.sp[t0 + q];
        cf[q] = d.coef[d.c16[t0 + q]];
    }
    const F en = E[(size_t)c * ncell + id];
    if constexpr (sizeof(F) == 4) {
        float w = 0.f;
#pragma unroll
        for (int q = 0; q < NP; ++q) w += pole_pair(sp[q], cf[q], en, mode);
#pragma unroll
        for (int q = 0; q < NP; ++q) d.sp[t0 + q] = sp[q];
        if (mode != 1) static_cast<float *>(d.hist)[mt.y] = w;
    } else {
        float2 sq[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) sq[q] = d.sq[t0 + q];
        float wr = 0.f, wi = 0.f;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            wr += pole_pair(sp[q], cf[q], en.r, mode);
            wi += pole_pair(sq[q], cf[q], en.i, mode);
        }
#pragma unroll
        for (int q = 0; q < NP; ++q) { d.sp[t0 + q] = sp[q]; d.sq[t0 + q] = sq[q]; }
        if (mode != 1) static_cast<cf2 *>(d.hist)[mt.y] = cf2{wr, wi};
    }
}

// Entry e: mode 0 = pre, 1 = post, 2 = fused post + pre.
template <typename F>
__device__ __forceinline__ void disp_entry(const DispDev &d, const F *__restrict__ E,
                                           size_t ncell, int ny, int nz, int mode, int e)
{
    switch (d.np) {
    case 1: disp_entry_np<F, 1>(d, E, ncell, mode, e); return;
    case 2: disp_entry_np<F, 2>(d, E, ncell, mode, e); return;
    case 3: disp_entry_np<F, 3>(d, E, ncell, mode, e); return;
    case 4: disp_entry_np<F, 4>(d, E, ncell, mode, e); return;
    default: break;
    }
    const int c = d.comp[e], id = d.cell[e];
    const F en = E[(size_t)c * ncell + id];
    if constexpr (sizeof(F) == 4) {
        float w = 0.f;
        for (int t = d.ofs[e]; t < d.ofs[e + 1]; ++t)
            w += pole_update(d.p_re, d.p_im, t, d.cidx[t], d, en, mode);
        if (mode != 1) static_cast<float *>(d.hist)[hist_index(d, c, id, ny, nz)] = w;
    } else {
        float wr = 0.f, wi = 0.f;
        for (int t = d.ofs[e]; t < d.ofs[e + 1]; ++t) {
            wr += pole_update(d.p_re, d.p_im, t, d.cidx[t], d, en.r, mode);
            wi += pole_update(d.q_re, d.q_im, t, d.cidx[t], d, en.i, mode);
        }
        if (mode != 1) static_cast<cf2 *>(d.hist)[hist_index(d, c, id, ny, nz)] = cf2{wr, wi};
    }
}

template <typename F>
__global__ void __launch_bounds__(256)
k_disp(DispDev d, const F *__restrict__ E, size_t ncell, int ny, int nz, int mode)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    disp_entry<F>(d, E, ncell, ny, nz, mode, e);
}

// ------------------------------------------------------------------------------------------
// Time-domain field sampling (reference sample_time_box, `openem/kernels/source_dft.cu:272-316`;
// complex sample_time_box_c, `openem/kernels/bloch.cu:278-300`).  At the end of step n (E^{n+1}
// and H^{n+1/2} in hand) shot 0 writes slot(m = n + 1) := E^m + H^{m-1/2} / 2 and shot 1 adds
// H^{m+1/2} / 2 into slot(m = n): H is reported as the average of the two half steps around
// m dt.  Layout (nc, slots, ni, nj, nk) of F at native Yee positions.  One thread per box cell.
// ------------------------------------------------------------------------------------------
struct TimeMonDev {
    int nc, comps[6];
    int o[3], b[3];
    int beg, end, iv, slots;
    int cells;
    void *buf;
};

template <typename F>
__global__ void __launch_bounds__(256)
k_time_sample(TimeMonDev d, const F *__restrict__ E, const F *__restrict__ H, size_t ncell,
              int ny, int nz, const int *__restrict__ stp, int shot)
{
    const int m = step_now(stp) + 1 - shot;
    const int off = m - d.beg;
    if (off < 0 || m >= d.end || off % d.iv) return;
    const int slot = off / d.iv;
    if (slot >= d.slots) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.cells) return;
    const int a = q / (d.b[1] * d.b[2]), r = q - a * d.b[1] * d.b[2];
    const int b = r / d.b[2], c = r - b * d.b[2];
    const size_t id = ((size_t)(d.o[0] + a) * ny + (d.o[1] + b)) * nz + (d.o[2] + c);
    F *out = static_cast<F *>(d.buf);
    for (int t = 0; t < d.nc; ++t) {
        const int comp = d.comps[t];
        if (shot == 1 && comp < 3) continue;
        const F v = comp < 3 ? E[(size_t)comp * ncell + id] : 0.5f * H[(size_t)(comp - 3) * ncell + id];
        F &o = out[((size_t)t * d.slots + slot) * d.cells + q];
        o = shot ? o + v : v;
    }
}

// ------------------------------------------------------------------------------------------
// Time-domain flux (reference _sample_flux_time, `openem/readout.py:20-73`, geometry
// `openem/flux.py:334-513`): one scalar per sampled step, sum_p w_p sum dS (E_u Hbar_v - E_v
// Hbar_u) over the colocated edge samples of the monitor plane(s).  At the end of step n the
// second shot of m = n adds the H^{n+1/2} half with the E planes kept by the previous step, and
// the first shot of m = n + 1 keeps the colocated planes of E^{n+1} and writes the H^{n+1/2}
// half.  Terms are formed in float32 like the fields, summed in float64: per-block partials
// (a fixed tree) and then one block summing the partials in block order, so the result does
// not depend on scheduling.
// ------------------------------------------------------------------------------------------
#define OEM_FT_BLK 256
struct FluxTimeDev {
    int axis, t1, t2, n1, n2, n_planes;
    const int *e1, *l1, *h1, *e2, *l2, *h2;     // edge sample and its lower / upper cell
    const float *w1, *w2;                       // centre -> edge weight of the lower cell
    const double *tr1, *tr2;                    // trapezoid weights
    double plane_w[2];
    int ke[2], kh[2];
    float wh_lo[2], wh_hi[2];
    double sign;
    int beg, end, iv, slots;
    float *ep;                                  // kept E planes: (plane, {u, v}, n1 * n2)
    double *partial;                            // (2, n_blocks): first / second shot
    double *buf;                                // (slots,)
    int n_blocks;
};

__device__ __forceinline__ int ft_slot(const FluxTimeDev &d, int m) {
    const int off = m - d.beg;
    if (off < 0 || m >= d.end || off % d.iv) return -1;
    const int s = off / d.iv;
    return s < d.slots ? s : -1;
}

__device__ __forceinline__ size_t ft_cell(const FluxTimeDev &d, int k, int i1, int i2, int ny,
                                          int nz) {
    int c[3];
    c[d.axis] = k;
    c[d.t1] = i1;
    c[d.t2] = i2;
    return ((size_t)c[0] * ny + c[1]) * nz + c[2];
}

// Sum of a double over the block (fixed tree), valid in thread 0.
__device__ __forceinline__ double ft_block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = OEM_FT_BLK / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    return sh[0];
}

template <typename F>
__global__ void __launch_bounds__(OEM_FT_BLK)
k_flux_time_partial(FluxTimeDev d, const F *__restrict__ E, const F *__restrict__ H,
                    size_t ncell, int ny, int nz, const int *__restrict__ stp)
{
    static_assert(sizeof(F) == 4, "flux-time monitors sample real fields");
    const int n = step_now(stp);
    const int s1 = ft_slot(d, n), s0 = ft_slot(d, n + 1);
    if (s1 < 0 && s0 < 0) return;                       // uniform over the grid
    __shared__ double sh[OEM_FT_BLK];
    const int np = d.n1 * d.n2;
    const int q = blockIdx.x * OEM_FT_BLK + threadIdx.x;
    double v0 = 0.0, v1 = 0.0;
    if (q < np) {
        const int a = q / d.n2, b = q - a * d.n2;
        const float w1 = d.w1[a], w2 = d.w2[b];
        const double dS = d.sign * d.tr1[a] * d.tr2[b];
        const float *Eu = E + (size_t)d.t1 * ncell, *Ev = E + (size_t)d.t2 * ncell;
        const float *Hu = H + (size_t)d.t1 * ncell, *Hv = H + (size_t)d.t2 * ncell;
        for (int p = 0; p < d.n_planes; ++p) {
            const int ke = d.ke[p], kh = d.kh[p];
            const float lo = d.wh_lo[p], hi = d.wh_hi[p];
            // H_u: on the edge along t1, centre -> edge along t2, coordinate weights along the normal
            const float hu_l = lo * Hu[ft_cell(d, kh, d.e1[a], d.l2[b], ny, nz)] +
                               hi * Hu[ft_cell(d, kh + 1, d.e1[a], d.l2[b], ny, nz)];
            const float hu_h = lo * Hu[ft_cell(d, kh, d.e1[a], d.h2[b], ny, nz)] +
                               hi * Hu[ft_cell(d, kh + 1, d.e1[a], d.h2[b], ny, nz)];
            const float hu = w2 * hu_l + (1.0f - w2) * hu_h;
            // H_v: centre -> edge along t1, on the edge along t2
            const float hv_l = lo * Hv[ft_cell(d, kh, d.l1[a], d.e2[b], ny, nz)] +
                               hi * Hv[ft_cell(d, kh + 1, d.l1[a], d.e2[b], ny, nz)];
            const float hv_h = lo * Hv[ft_cell(d, kh, d.h1[a], d.e2[b], ny, nz)] +
                               hi * Hv[ft_cell(d, kh + 1, d.h1[a], d.e2[b], ny, nz)];
            const float hv = w1 * hv_l + (1.0f - w1) * hv_h;
            float *keep = d.ep + (size_t)(2 * p) * np;
            const double wp = 0.5 * d.plane_w[p] * dS;
            if (s1 >= 0)
                v1 += wp * ((double)(keep[q] * hv) - (double)(keep[np + q] * hu));
            if (s0 >= 0) {
                const float eu = w1 * Eu[ft_cell(d, ke, d.l1[a], d.e2[b], ny, nz)] +
                                 (1.0f - w1) * Eu[ft_cell(d, ke, d.h1[a], d.e2[b], ny, nz)];
                const float ev = w2 * Ev[ft_cell(d, ke, d.e1[a], d.l2[b], ny, nz)] +
                                 (1.0f - w2) * Ev[ft_cell(d, ke, d.e1[a], d.h2[b], ny, nz)];
                v0 += wp * ((double)(eu * hv) - (double)(ev * hu));
                keep[q] = eu;
                keep[np + q] = ev;
            }
        }
    }
    const double t0 = ft_block_sum(v0, sh);
    __syncthreads();
    const double t1 = ft_block_sum(v1, sh);
    if (threadIdx.x == 0) {
        d.partial[blockIdx.x] = t0;
        d.partial[d.n_blocks + blockIdx.x] = t1;
    }
}

template <typename F>
__global__ void __launch_bounds__(OEM_FT_BLK)
k_flux_time_sum(FluxTimeDev d, const int *__restrict__ stp)
{
    const int n = step_now(stp);
    const int s1 = ft_slot(d, n), s0 = ft_slot(d, n + 1);
    if (s1 < 0 && s0 < 0) return;
    __shared__ double sh[OEM_FT_BLK];
    double a0 = 0.0, a1 = 0.0;
    for (int b = threadIdx.x; b < d.n_blocks; b += OEM_FT_BLK) {
        a0 += d.partial[b];
        a1 += d.partial[d.n_blocks + b];
    }
    const double t0 = ft_block_sum(a0, sh);
    __syncthreads();
    const double t1 = ft_block_sum(a1, sh);
    if (threadIdx.x == 0) {
        if (s1 >= 0) d.buf[s1] += t1;
        if (s0 >= 0) d.buf[s0] = t0;
    }
}

// ------------------------------------------------------------------------------------------
// Running DFT, batched: every step a snapshot of the monitored values goes into a ring of T
// rows; every T steps one flush replays them in step order against the float64 phase table and
// adds into the float64 accumulators (reference accumulate_dft / accumulate_dft_box and the
// batched dft_snap_* / dft_flush_*, `openem/kernels/source_dft.cu:216-271`, `:371-430`,
// `:453-779`; there the partial sums are float32, here they are float64).  Complex fields
// accumulate (u_r + i u_i) e^{i w t}, the reference's complex DFT (`bloch.cu:389-481`).
// ------------------------------------------------------------------------------------------
struct MonDev {
    int kind;                // 0 plane, 1 box
    int axis, km, n1, n2;    // plane
    int o[3], b[3];          // box
    int ncomp, npts, nf, T;
    void *snap;              // T * ncomp * npts elements of F
    double *re, *im;         // ncomp * nf * npts
    const double *phase;     // n_steps * nf * 4
};

// Point q of a monitor into ring row `slot`.
template <typename F>
__device__ __forceinline__ void snap_point(const MonDev &m, const F *__restrict__ E,
                                           const F *__restrict__ H, size_t ncell, int ny, int nz,
                                           int slot, int q)
{
    F *row = static_cast<F *>(m.snap) + (size_t)slot * m.ncomp * m.npts;
    if (m.kind == 0) {
        const int a = q / m.n2, b = q - a * m.n2;
        int c[3];
        const int t1 = m.axis == 0 ? 1 : 0, t2 = m.axis == 2 ? 1 : 2;
        c[m.axis] = m.km; c[t1] = a; c[t2] = b;
        const size_t id = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        c[m.axis] = m.km - 1;
        const size_t idm = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        row[q] = E[(size_t)t1 * ncell + id];
        row[m.npts + q] = E[(size_t)t2 * ncell + id];
        row[2 * m.npts + q] = 0.5f * (H[(size_t)t1 * ncell + id] + H[(size_t)t1 * ncell + idm]);
        row[3 * m.npts + q] = 0.5f * (H[(size_t)t2 * ncell + id] + H[(size_t)t2 * ncell + idm]);
    } else {
        const int bjk = m.b[1] * m.b[2];
        const int a = q / bjk, r = q - a * bjk, b = r / m.b[2], cc = r - b * m.b[2];
        const size_t id = ((size_t)(m.o[0] + a) * ny + (m.o[1] + b)) * nz + (m.o[2] + cc);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            row[(size_t)c * m.npts + q] = E[(size_t)c * ncell + id];
            row[(size_t)(3 + c) * m.npts + q] = H[(size_t)c * ncell + id];
        }
    }
}

template <typename F>
__global__ void __launch_bounds__(256)
k_dft_snap(MonDev m, const F *__restrict__ E, const F *__restrict__ H, size_t ncell, int ny,
           int nz, const int *__restrict__ stp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.npts) return;
    snap_point<F>(m, E, H, ncell, ny, nz, step_now(stp) % m.T, q);
}

// The end of a step as one single-block launch, where three small launches would otherwise
// close it: the E-side point lists in step order (each list after the one before it), then the
// monitor points those lists corrected, snapped again, then the step counter.  Used when every
// list fits the block and nothing else reads the step counter after the E-side points.
#define OEM_TAIL_LISTS 4
#define OEM_TAIL_THREADS 1024
struct TailDev {
    int n_pts, n_snap;
    PointsDev pts[OEM_TAIL_LISTS];
    MonDev mon[OEM_TAIL_LISTS];
    const int *qs[OEM_TAIL_LISTS];
    int nq[OEM_TAIL_LISTS];
};

template <typename F>
__global__ void __launch_bounds__(OEM_TAIL_THREADS)
k_step_tail(F *__restrict__ E, const F *__restrict__ H, TailDev t, MatDev mat, float ch,
            size_t ncell, int ny, int nz, int *__restrict__ stp)
{
    const int n = step_now(stp);
    const int tid = threadIdx.x;
    for (int p = 0; p < t.n_pts; ++p) {
        if (tid < t.pts[p].n_groups) point_group<F>(E, t.pts[p], mat, ch, ncell, n, tid);
        __syncthreads();
    }
    for (int m = 0; m < t.n_snap; ++m)
        if (tid < t.nq[m]) snap_point<F>(t.mon[m], E, H, ncell, ny, nz, n % t.mon[m].T, t.qs[m][tid]);
    __syncthreads();
    if (tid == 0) *stp = n + 1;
}

// The points `qs[0..nq)` of a monitor whose other points the fused step snapped.
template <typename F>
__global__ void __launch_bounds__(256)
k_dft_snap_list(MonDev m, const F *__restrict__ E, const F *__restrict__ H, size_t ncell, int ny,
                int nz, const int *__restrict__ stp, const int *__restrict__ qs, int nq)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    snap_point<F>(m, E, H, ncell, ny, nz, step_now(stp) % m.T, qs[t]);
}

__device__ __forceinline__ void dft_acc(double &sr, double &si, float v, double c, double s) {
    sr += (double)v * c;
    si += (double)v * s;
}
__device__ __forceinline__ void dft_acc(double &sr, double &si, cf2 v, double c, double s) {
    sr += (double)v.r * c - (double)v.i * s;
    si += (double)v.r * s + (double)v.i * c;
}

// Replay the snapshots of steps first..last (in step order) for point q at frequency f; NC
// components of which the first NC/2 are E (whole-step phase).
template <typename F, int NC>
__device__ __forceinline__ void flush_point(const MonDev &m, int first, int last, int f, int q)
{
    double sre[NC], sim[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) { sre[c] = 0.0; sim[c] = 0.0; }
    for (int s = first; s <= last; ++s) {
        const F *row = static_cast<const F *>(m.snap) + (size_t)(s % m.T) * NC * m.npts;
        const double *ph = m.phase + ((size_t)s * m.nf + f) * 4;
        const double ce = ph[0], se = ph[1], chh = ph[2], shh = ph[3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const F v = row[(size_t)c * m.npts + q];
            dft_acc(sre[c], sim[c], v, c < NC / 2 ? ce : chh, c < NC / 2 ? se : shh);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const size_t o = ((size_t)c * m.nf + f) * m.npts + q;
        m.re[o] += sre[c];
        m.im[o] += sim[c];
    }
}

// Replay of a single-scene monitor: a thread per (point, component) and group of up to
// OEM_FLUSH_FG frequencies, so each snapshot value is read once per group, not once per
// frequency (the phase rows are the same for the whole block and come through scalar loads).
// Each (point, component, frequency) sum is formed in the order flush_point forms it, so both
// give the same bits.  grid: x over points, y over (component, frequency group).  Launched after
// the step counter advanced past the last snapshot: the cnt snapshots ending at step *stp - 1
// are replayed.
#define OEM_FLUSH_FG 8
template <typename F>
__global__ void __launch_bounds__(256)
k_dft_flush(MonDev m, int cnt, const int *__restrict__ stp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int ngrp = (m.nf + OEM_FLUSH_FG - 1) / OEM_FLUSH_FG;
    const int c = blockIdx.y / ngrp, f0 = (blockIdx.y - c * ngrp) * OEM_FLUSH_FG;
    if (q >= m.npts || c >= m.ncomp) return;
    const int nfl = min(OEM_FLUSH_FG, m.nf - f0);
    const int last = step_now(stp) - 1, first = last - cnt + 1;
    const int po = c < m.ncomp / 2 ? 0 : 2;        // E: whole-step phase, H: half-step phase
    double sre[OEM_FLUSH_FG], sim[OEM_FLUSH_FG];
#pragma unroll
    for (int u = 0; u < OEM_FLUSH_FG; ++u) { sre[u] = 0.0; sim[u] = 0.0; }
    for (int s = first; s <= last; ++s) {
        const F v = static_cast<const F *>(m.snap)[((size_t)(s % m.T) * m.ncomp + c) * m.npts + q];
        const double *ph = m.phase + ((size_t)s * m.nf + f0) * 4 + po;
#pragma unroll
        for (int u = 0; u < OEM_FLUSH_FG; ++u)
            if (u < nfl) dft_acc(sre[u], sim[u], v, ph[4 * u], ph[4 * u + 1]);
    }
#pragma unroll
    for (int u = 0; u < OEM_FLUSH_FG; ++u)
        if (u < nfl) {
            const size_t o = ((size_t)c * m.nf + f0 + u) * m.npts + q;
            m.re[o] += sre[u];
            m.im[o] += sim[u];
        }
}
