// Discrete adjoint of the time step (real or complex Bloch fields, with or without ADE poles):
// checkpoints, recomputation of a segment with E recorded at the design cells, and the backward
// sweep.
//
// The forward step is linear in (E, H, psi_H, psi_E) plus source terms; with
//   H' = H - ch (s1 - s2),  s = (1/kappa + a) D_fwd E + b psi_H,  psi_H' = b psi_H + a D_fwd E
//   E' = pec (ca E + cb (s1 - s2)),  s = (1/kappa + a) D_bwd H' + b psi_E,  psi_E' = ...
// the backward sweep applies, per step n from the last: the objective's gradient with respect to
// E^{n+1} and H^{n+1/2} (monitor samples), the eps gradient lambda_E . (ca' E^n + cb' S^n), the
// transpose of the E update (gather from the forward neighbour along each axis, then the
// cell-local scaling), and the transpose of the H update (gather from the backward neighbour).
// Each transpose is one kernel that reads the old adjoint fields and writes the new cell-local
// ones (lambda_E, the psi adjoints) to a second buffer, so no thread reads a value another
// overwrites; the buffers swap after each kernel.
//
// ADE poles (per entry: component c at a cell, complex pole states P_p driven by the real E)
// add two cell-local transposes per step, one kernel over the entries before the eps gradient
// (k_adj_disp): the post update P^{n+1} = P_mid + B E^{n+1} gives lambda_E^{n+1} the term
// sum_p Re(conj(B_p) lambda_P_p); the pre update dP = (A-1) P + B E^n, hist = sum 2 Re dP, fed
// into E^{n+1} as -cw cb pec hist, gives lambda_dP = lambda_P + 2 (-cw cb pec lambda_E^{n+1}),
// lambda_P <- lambda_P + conj(A-1) lambda_dP and lambda_E^n the term sum_p Re(conj(B_p)
// lambda_dP_p).  That last term is kept per entry until the H transpose has written
// lambda_E^n, and the next step's k_adj_disp adds it first.
//
// Complex (Bloch) fields: every coefficient is real except the wrap masks of the Bloch axes
// (the phase a neighbour across the boundary carries), so the step is complex-linear and its
// adjoint, with lambda = dJ/d Re + i dJ/d Im, is the conjugate transpose: the same kernels on
// complex values with the masks conjugated.  The pole state driven by Im E has its own adjoint
// (lQ), fed by the imaginary parts as lP is by the real ones; the eps gradient is
// Re(conj(lambda_E) (ca' E^n + cb' S^n)).
//
// Each term of an update: (target component c, source component b, derivative axis a, sign,
// psi slot).  psi slot = off[2 a + m] of the slab layout (x: Hy/Ey 0, Hz/Ez 1; y: Hx/Ex 0,
// Hz/Ez 1; z: Hx/Ex 0, Hy/Ey 1).

#include "oem_internal.h"

#include <unordered_set>

// the transposed sweeps: a (j, k) column per thread from a block of 256, marching `chunk` x
// planes per block (grid z; OEM_ADJ_CHUNK by default, the environment's OEM_ADJ_CHUNK if set)
#define OEM_ADJ_TK 64
#ifndef OEM_ADJ_TJ
#define OEM_ADJ_TJ 4
#endif
// 8 planes measured best on the 400 x 400 x 48 bend (s per design: 4 3.04-3.06, 6 3.01, 8
// 3.02-3.09, 12 3.08, 16 3.11-3.18, 24 3.10, 32 3.17)
#define OEM_ADJ_CHUNK 8

namespace {

struct AdjDev {
    float *lE, *lH;              // 3 * ncell each
    float *lpe, *lph;            // psi adjoints, slab layout of PsiDev
    float *lpe_n, *lph_n;        // where the transposes write the new lpe, lph (ping-pong)
    PsiDev psl;                  // offsets (pointers unused)
    GridDev g;
    MatDev mat;
    float ch;
    size_t ncell;
    int chunk;                   // x planes per block of the transposes
};

// E transpose, part 1: lH_b += D_bwd^T q_{c,a} for the two terms with source b.
// D_bwd f[p] = (f[p] - f[prv p] mpv[p]) idd[p]  =>  (D_bwd^T q)[p] = q[p] idd[p]
//   - sum over p' with prv[p'] == p of q[p'] mpv[p'] idd[p']  (p' = nxt p, or p itself at a
//   clamped end)
// A thread keeps one (j, k) column and marches a chunk of x planes downwards, so the x term of
// plane i + 1 is carried from the previous iteration; the y and z neighbours' terms are
// recomputed from their cells.  Terms (component -> source, axis): t0 c0->b2 y, t1 c0->b1 z, t2 c1->b0 z,
// t3 c1->b2 x, t4 c2->b1 x, t5 c2->b0 y.

// One axis' table values at coordinate p: the CPML (a, 1/kappa) of one side, the slab index,
// 1/dl and the mask of the difference that the transposed term uses.
struct AxAt {
    float a, b, ik, idl, m, mi, pec;    // mi: imaginary part of the mask (Bloch wraps)
    int s;
};
__device__ __forceinline__ int mask_im(int fm) { return fm == AX_MNX ? AX_MNXI : AX_MPVI; }
__device__ __forceinline__ AxAt ax_at(const AxisDev &A, int n, int p, int fa, int fk, int fidl,
                                      int fm) {
    AxAt r;
    r.s = slab_of(p, n, A.lo, A.hi);
    r.a = axv(A, p, fa); r.ik = axv(A, p, fk);
    r.b = axv(A, p, fa == AX_AE ? AX_BE : AX_BH);
    r.idl = axv(A, p, fidl); r.m = axv(A, p, fm); r.mi = axv(A, p, mask_im(fm));
    r.pec = axv(A, p, AX_PEC);
    return r;
}
// ax_at of a wave-uniform plane p, through scalar loads
__device__ __forceinline__ AxAt ax_at_row(const AxisDev &A, int n, int p, int fa, int fk, int fidl,
                                          int fm) {
    const ctab_t t = ctab_row(A, p);
    AxAt r;
    r.s = slab_of(p, n, A.lo, A.hi);
    r.a = t[fa]; r.ik = t[fk];
    r.b = t[fa == AX_AE ? AX_BE : AX_BH];
    r.idl = t[fidl]; r.m = t[fm]; r.mi = t[mask_im(fm)];
    r.pec = t[AX_PEC];
    return r;
}

// q of one transposed term: (1/kappa + a) sgn x + a psi in the slab, sgn x outside
template <typename F>
__device__ __forceinline__ F q_of(F sgn_x, const AxAt &t, float sgn, F x, F psi) {
    return t.s >= 0 ? (t.ik + t.a) * sgn * x + t.a * psi : sgn_x;
}

__device__ __forceinline__ void stage_lut_adj(float4 *s_lut, const MatDev &mat) {
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const float4 *src = reinterpret_cast<const float4 *>(mat.lut);
    for (int q = tid; q < 2 * mat.n_classes; q += blockDim.x * blockDim.y) s_lut[q] = src[q];
}

template <typename F>
__global__ void __launch_bounds__(OEM_ADJ_TK * OEM_ADJ_TJ) k_adj_e1(AdjDev d)
{
    // F: float, or cf2 on complex (Bloch) fields; EB bytes per element, SH the shift from
    // an element's byte offset to its class index's
    typedef FT<F> T;
    constexpr uint32_t EB = T::B, SH = EB == 4 ? 1 : 2;
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    const MatDev &mat = d.mat;
    const bool lds = lut_in_lds(mat);
    if (lds) {
        stage_lut_adj(s_lut, mat);
        __syncthreads();
    }
    const GridDev &g = d.g;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // one (j, k) column per thread, numbered along the plane: no idle lanes when nz is not a
    // multiple of 64 (a 48-cell-thick slab leaves a quarter of (k, j) tiles empty).  Threads
    // past the last column take the last one and store nothing: every thread reaches the
    // per-plane barrier.
    const int tid = threadIdx.y * OEM_ADJ_TK + threadIdx.x;
    const int base = blockIdx.x * (OEM_ADJ_TK * OEM_ADJ_TJ);
    const bool act = base + tid < ny * nz;
    const int q = act ? base + tid : ny * nz - 1;
    const int j = q / nz, k = q - j * nz;
    const int i0 = blockIdx.z * d.chunk, i1 = min(i0 + d.chunk, nx);
    const uint32_t nc = (uint32_t)d.ncell, plane = (uint32_t)(ny * nz);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    const uint32_t *off = d.psl.off;
    // buffer resources (the x march's per-plane offsets go in SGPRs): the three components of
    // lambda_E (old, new) and lambda_H at c * cB, the psi adjoints (old, new), the classes
    const uint32_t cB = nc * EB;
    const rsrc_t rE = mk_rsrc(d.lE, 3u * cB);
    const rsrc_t rH = mk_rsrc(d.lH, 3u * cB);
    const rsrc_t rP = mk_rsrc(d.lpe, d.psl.total * EB), rPn = mk_rsrc(d.lpe_n, d.psl.total * EB);
    const rsrc_t rC = mk_rsrc(mat.cls, nc * 2u);
    // the y and z terms every thread computes for its own cell, published per plane so the
    // threads of the backward neighbours (j - 1, k - 1 in the interior) read them instead of
    // recomputing them from the neighbour's cell (double-buffered by plane parity: one barrier
    // per plane)
    __shared__ F s_q[2][4][OEM_ADJ_TK * OEM_ADJ_TJ];
    // y and z at the cell and at the forward neighbours Y.nxt[j], Z.nxt[k]
    const AxAt yc = ax_at(Y, ny, j, AX_AE, AX_KE, AX_IDLD, AX_MPV);
    const AxAt zc = ax_at(Z, nz, k, AX_AE, AX_KE, AX_IDLD, AX_MPV);
    const int jn = Y.nxt[j], kn = Z.nxt[k];
    const bool y_lo = Y.prv[j] == j, z_lo = Z.prv[k] == k;
    const bool y_nb = jn != j && Y.prv[jn] == j, z_nb = kn != k && Z.prv[kn] == k;
    const AxAt yn = ax_at(Y, ny, jn, AX_AE, AX_KE, AX_IDLD, AX_MPV);
    const AxAt zn = ax_at(Z, nz, kn, AX_AE, AX_KE, AX_IDLD, AX_MPV);
    // the neighbours' threads in this block, if they are in it
    const int ly = jn * nz + k - base, lz = j * nz + kn - base;
    const bool y_lds = y_nb && ly >= 0 && ly < OEM_ADJ_TK * OEM_ADJ_TJ;
    const bool z_lds = z_nb && lz >= 0 && lz < OEM_ADJ_TK * OEM_ADJ_TJ;
    // per-thread byte offsets within a plane: the cell, its y and z neighbours, their psi rows
    const uint32_t vo = (uint32_t)(j * nz + k) * EB, vyn = (uint32_t)(jn * nz + k) * EB,
                   vzn = (uint32_t)(j * nz + kn) * EB;
    const uint32_t vpy = (uint32_t)(max(yc.s, 0) * nz + k) * EB;
    const uint32_t vpyn = (uint32_t)(max(yn.s, 0) * nz + k) * EB;
    const uint32_t vpz = (uint32_t)(j * wz + max(zc.s, 0)) * EB;
    const uint32_t vpzn = (uint32_t)(j * wz + max(zn.s, 0)) * EB;
    int buf = 0;
    // the class of the cell at byte offset v of plane offset sp (class LUT materials), and the
    // cell's cb / ca from its class or from the dense arrays (flat cell id)
    auto cls_at = [&](uint32_t v, uint32_t sp) -> uint32_t {
        return mat.mode == 0 ? bld_u16(rC, v >> SH, sp >> SH) : 0u;
    };
    auto cbs = [&](uint32_t c, uint32_t id, float &bx, float &by, float &bz) {
        if (mat.mode == 0) {
            if (lds) {
                const float4 r0 = s_lut[2 * c], r1 = s_lut[2 * c + 1];
                bx = r0.y; by = r0.w; bz = r1.y;
            } else {
                bx = mat.lut[8 * (size_t)c + 1]; by = mat.lut[8 * (size_t)c + 3]; bz = mat.lut[8 * (size_t)c + 5];
            }
        } else {
            bx = mat.cb[id]; by = mat.cb[nc + id]; bz = mat.cb[2 * nc + id];
        }
    };

    // the x terms (t3: c1, sgn -1, slot 0; t4: c2, sgn +1, slot 1) of a cell on plane i
    auto qx = [&](int i, F &q3, F &q4) {
        const uint32_t sp = (uint32_t)i * plane * EB;
        const AxAt xc = ax_at_row(X, nx, i, AX_AE, AX_KE, AX_IDLD, AX_MPV);
        float bx, by, bz;
        cbs(cls_at(vo, sp), (uint32_t)i * plane + (uint32_t)(j * nz + k), bx, by, bz);
        const F nu1 = by * (zc.pec * xc.pec) * T::ld(rE, vo, sp + cB);
        const F nu2 = bz * (xc.pec * yc.pec) * T::ld(rE, vo, sp + 2u * cB);
        F p0 = T::zero(), p1 = T::zero();
        if (xc.s >= 0) {
            const uint32_t sx = (uint32_t)xc.s * plane * EB;
            p0 = T::ld(rP, vo, off[0] * EB + sx);
            p1 = T::ld(rP, vo, off[1] * EB + sx);
        }
        q3 = q_of(-1.f * nu1, xc, -1.f, nu1, p0);
        q4 = q_of(1.f * nu2, xc, 1.f, nu2, p1);
        (void)bx;
    };
    // carried x terms of the plane above the current one
    F q3n = T::zero(), q4n = T::zero();
    bool x_nb_top = false;
    {
        int inb = i1;
        if (i1 < nx) {
            x_nb_top = true;
        } else {
            inb = X.nxt[nx - 1];
            x_nb_top = inb != nx - 1 && X.prv[inb] == nx - 1;
        }
        if (x_nb_top) qx(inb, q3n, q4n);
    }
    for (int i = i1 - 1; i >= i0; --i) {
        const uint32_t id = (uint32_t)i * plane + (uint32_t)(j * nz + k);
        const uint32_t sp = (uint32_t)i * plane * EB;
        const uint32_t sy = (uint32_t)i * (uint32_t)(wy * nz) * EB;
        const uint32_t sz = (uint32_t)i * (uint32_t)(ny * wz) * EB;
        const AxAt xc = ax_at_row(X, nx, i, AX_AE, AX_KE, AX_IDLD, AX_MPV);
        const uint32_t sx = (uint32_t)max(xc.s, 0) * plane * EB;
        const uint32_t cls = cls_at(vo, sp);
        const F l0 = T::ld(rE, vo, sp), l1 = T::ld(rE, vo, sp + cB),
                    l2 = T::ld(rE, vo, sp + 2u * cB);
        F py0 = T::zero(), py1 = T::zero(), pz0 = T::zero(), pz1 = T::zero(), px0 = T::zero(), px1 = T::zero();
        if (yc.s >= 0) { py0 = T::ld(rP, vpy, off[2] * EB + sy); py1 = T::ld(rP, vpy, off[3] * EB + sy); }
        if (zc.s >= 0) { pz0 = T::ld(rP, vpz, off[4] * EB + sz); pz1 = T::ld(rP, vpz, off[5] * EB + sz); }
        if (xc.s >= 0) { px0 = T::ld(rP, vo, off[0] * EB + sx); px1 = T::ld(rP, vo, off[1] * EB + sx); }
        // every global read of the plane goes out before the barrier (the cell's lambda_H, and
        // the neighbours' terms recomputed from their cells where the neighbour's thread is not
        // in the block), so nothing after it waits on a load: a wait there would also wait for
        // the stores in flight (loads and stores share one in-order counter)
        F hl0 = T::zero(), hl1 = T::zero(), hl2 = T::zero();
        if (act) {
            hl0 = T::ld(rH, vo, sp);
            hl1 = T::ld(rH, vo, sp + cB);
            hl2 = T::ld(rH, vo, sp + 2u * cB);
        }
        // neighbours' terms: y (t0, t5) at Y.nxt[j], z (t1, t2) at Z.nxt[k], x carried; from
        // the neighbour's thread when it is in the block (after the barrier), else recomputed
        // here with the same expressions
        F q0n = T::zero(), q5n = T::zero(), q1n = T::zero(), q2n = T::zero();
        if (!y_lds && y_nb) {
            float b0, b1, b2;
            cbs(cls_at(vyn, sp), (uint32_t)i * plane + (uint32_t)(jn * nz + k), b0, b1, b2);
            const F n0 = b0 * (yn.pec * zc.pec) * T::ld(rE, vyn, sp);
            const F n2 = b2 * (xc.pec * yn.pec) * T::ld(rE, vyn, sp + 2u * cB);
            F a0 = T::zero(), a1 = T::zero();
            if (yn.s >= 0) { a0 = T::ld(rP, vpyn, off[2] * EB + sy); a1 = T::ld(rP, vpyn, off[3] * EB + sy); }
            q0n = q_of(1.f * n0, yn, 1.f, n0, a0);
            q5n = q_of(-1.f * n2, yn, -1.f, n2, a1);
            (void)b1;
        }
        if (!z_lds && z_nb) {
            float b0, b1, b2;
            cbs(cls_at(vzn, sp), (uint32_t)i * plane + (uint32_t)(j * nz + kn), b0, b1, b2);
            const F n0 = b0 * (yc.pec * zn.pec) * T::ld(rE, vzn, sp);
            const F n1 = b1 * (zn.pec * xc.pec) * T::ld(rE, vzn, sp + cB);
            F a0 = T::zero(), a1 = T::zero();
            if (zn.s >= 0) { a0 = T::ld(rP, vpzn, off[4] * EB + sz); a1 = T::ld(rP, vpzn, off[5] * EB + sz); }
            q1n = q_of(-1.f * n0, zn, -1.f, n0, a0);
            q2n = q_of(1.f * n1, zn, 1.f, n1, a1);
            (void)b2;
        }
        float bx, by, bz;
        cbs(cls, id, bx, by, bz);
        // nu_c = cb_c pec_c lE_c
        const F nu0 = bx * (yc.pec * zc.pec) * l0;
        const F nu1 = by * (zc.pec * xc.pec) * l1;
        const F nu2 = bz * (xc.pec * yc.pec) * l2;
        const F q0 = q_of(1.f * nu0, yc, 1.f, nu0, py0);      // t0: c0, y, slot 2
        const F q5 = q_of(-1.f * nu2, yc, -1.f, nu2, py1);    // t5: c2, y, slot 3
        const F q1 = q_of(-1.f * nu0, zc, -1.f, nu0, pz0);    // t1: c0, z, slot 4
        const F q2 = q_of(1.f * nu1, zc, 1.f, nu1, pz1);      // t2: c1, z, slot 5
        const F q3 = q_of(-1.f * nu1, xc, -1.f, nu1, px0);    // t3: c1, x, slot 0
        const F q4 = q_of(1.f * nu2, xc, 1.f, nu2, px1);      // t4: c2, x, slot 1
        s_q[buf][0][tid] = q0;
        s_q[buf][1][tid] = q5;
        s_q[buf][2][tid] = q1;
        s_q[buf][3][tid] = q2;
        __syncthreads();
        if (y_lds) {
            q0n = s_q[buf][0][ly];
            q5n = s_q[buf][1][ly];
        }
        if (z_lds) {
            q1n = s_q[buf][2][lz];
            q2n = s_q[buf][3][lz];
        }
        const bool x_lo = cidx_at(X.prv, i) == i;
        const bool x_nb = i + 1 < i1 || x_nb_top;
        // the neighbour's mask and 1/dl along x: plane i + 1, or the wrap plane at the top
        const int ixn = i + 1 < nx ? i + 1 : cidx_at(X.nxt, nx - 1);
        const float mxn = ctab_row(X, ixn)[AX_MPV], mxni = ctab_row(X, ixn)[AX_MPVI],
                    ixdn = ctab_row(X, ixn)[AX_IDLD];
        // conj(mask): the transpose of a complex-linear step is its conjugate transpose
        auto term = [&](F q, const AxAt &c, bool lo, bool nb, F qn, float mn, float mni,
                        float idn) {
            F v = q * c.idl;
            if (lo) v -= T::mmul(q, c.m, -c.mi) * c.idl;
            if (nb) v -= T::mmul(qn, mn, -mni) * idn;
            return v;
        };
        const F v0 = term(q0, yc, y_lo, y_nb, q0n, yn.m, yn.mi, yn.idl);
        const F v1 = term(q1, zc, z_lo, z_nb, q1n, zn.m, zn.mi, zn.idl);
        const F v2 = term(q2, zc, z_lo, z_nb, q2n, zn.m, zn.mi, zn.idl);
        const F v3 = term(q3, xc, x_lo, x_nb, q3n, mxn, mxni, ixdn);
        const F v4 = term(q4, xc, x_lo, x_nb, q4n, mxn, mxni, ixdn);
        const F v5 = term(q5, yc, y_lo, y_nb, q5n, yn.m, yn.mi, yn.idl);
        buf ^= 1;
        q3n = q3;
        q4n = q4;
        if (!act) continue;
        F acc0 = T::zero(), acc1 = T::zero(), acc2 = T::zero();
        acc2 += v0; acc1 += v1; acc0 += v2; acc2 += v3; acc1 += v4; acc0 += v5;
        T::st(rH, hl0 + acc0, vo, sp);
        T::st(rH, hl1 + acc1, vo, sp + cB);
        T::st(rH, hl2 + acc2, vo, sp + 2u * cB);
        // the cell-local part of the psi adjoints, into the other buffer (neighbours read the
        // old values): lpe <- b (sgn nu + lpe) per slab term.  The cell-local part of lambda_E
        // (ca pec lambda_E) is left to the H transpose, which rewrites lambda_E in place
        if (xc.s >= 0) {
            T::st(rPn, xc.b * (-1.f * nu1 + px0), vo, off[0] * EB + sx);
            T::st(rPn, xc.b * (1.f * nu2 + px1), vo, off[1] * EB + sx);
        }
        if (yc.s >= 0) {
            T::st(rPn, yc.b * (1.f * nu0 + py0), vpy, off[2] * EB + sy);
            T::st(rPn, yc.b * (-1.f * nu2 + py1), vpy, off[3] * EB + sy);
        }
        if (zc.s >= 0) {
            T::st(rPn, zc.b * (-1.f * nu0 + pz0), vpz, off[4] * EB + sz);
            T::st(rPn, zc.b * (1.f * nu1 + pz1), vpz, off[5] * EB + sz);
        }
    }
}

// H transpose, part 1: lE_b += D_fwd^T q_{c,a}.
// D_fwd f[p] = (f[nxt p] mnx[p] - f[p]) idp[p]  =>  (D_fwd^T q)[p] = -q[p] idp[p]
//   + sum over p' with nxt[p'] == p of q[p'] mnx[p'] idp[p']  (p' = prv p, or p itself at a
//   clamped end)
// The march runs upwards in x, carrying the x terms of plane i - 1; the y and z neighbours
// (Y.prv[j], Z.prv[k]) are recomputed.  q = (1/kappa + a) f + a psi_H, f = -ch sgn lH_c.
template <typename F>
__global__ void __launch_bounds__(OEM_ADJ_TK * OEM_ADJ_TJ) k_adj_h1(AdjDev d)
{
    // F: float, or cf2 on complex (Bloch) fields; EB bytes per element, SH the shift from
    // an element's byte offset to its class index's
    typedef FT<F> T;
    constexpr uint32_t EB = T::B, SH = EB == 4 ? 1 : 2;
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    const MatDev &mat = d.mat;
    const bool lds = lut_in_lds(mat);
    if (lds) {
        stage_lut_adj(s_lut, mat);
        __syncthreads();
    }
    const GridDev &g = d.g;
    const AxisDev &X = g.ax[0], &Y = g.ax[1], &Z = g.ax[2];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // one (j, k) column per thread, numbered along the plane (threads past the last column take
    // the last one and store nothing, so every thread reaches the per-plane barrier)
    const int tid = threadIdx.y * OEM_ADJ_TK + threadIdx.x;
    const int base = blockIdx.x * (OEM_ADJ_TK * OEM_ADJ_TJ);
    const bool act = base + tid < ny * nz;
    const int q = act ? base + tid : ny * nz - 1;
    const int j = q / nz, k = q - j * nz;
    const int i0 = blockIdx.z * d.chunk, i1 = min(i0 + d.chunk, nx);
    const uint32_t nc = (uint32_t)d.ncell, plane = (uint32_t)(ny * nz);
    const int wy = Y.lo + Y.hi, wz = Z.lo + Z.hi;
    const uint32_t *off = d.psl.off;
    const float mch = -d.ch;
    // buffer resources: lambda_H, lambda_E (three components at c * cB), the psi adjoints
    const uint32_t cB = nc * EB;
    const rsrc_t rH = mk_rsrc(d.lH, 3u * cB), rE = mk_rsrc(d.lE, 3u * cB);
    const rsrc_t rP = mk_rsrc(d.lph, d.psl.total * EB), rPn = mk_rsrc(d.lph_n, d.psl.total * EB);
    const rsrc_t rC = mk_rsrc(mat.cls, nc * 2u);
    // ca of the three components of a cell, from its class or the dense arrays (flat id)
    auto cas = [&](uint32_t v, uint32_t sp, uint32_t id, float &ax, float &ay, float &az) {
        if (mat.mode == 0) {
            const uint32_t c = bld_u16(rC, v >> SH, sp >> SH);
            if (lds) {
                const float4 r0 = s_lut[2 * c], r1 = s_lut[2 * c + 1];
                ax = r0.x; ay = r0.z; az = r1.x;
            } else {
                ax = mat.lut[8 * (size_t)c]; ay = mat.lut[8 * (size_t)c + 2]; az = mat.lut[8 * (size_t)c + 4];
            }
        } else {
            ax = mat.ca[id]; ay = mat.ca[nc + id]; az = mat.ca[2 * nc + id];
        }
    };
    // each thread's y and z terms, published per plane for the forward neighbours' threads
    __shared__ F s_q[2][4][OEM_ADJ_TK * OEM_ADJ_TJ];
    const AxAt yc = ax_at(Y, ny, j, AX_AH, AX_KH, AX_IDLP, AX_MNX);
    const AxAt zc = ax_at(Z, nz, k, AX_AH, AX_KH, AX_IDLP, AX_MNX);
    const int jp = Y.prv[j], kp = Z.prv[k];
    const bool y_hi = Y.nxt[j] == j, z_hi = Z.nxt[k] == k;
    const bool y_nb = jp != j && Y.nxt[jp] == j, z_nb = kp != k && Z.nxt[kp] == k;
    const AxAt yp = ax_at(Y, ny, jp, AX_AH, AX_KH, AX_IDLP, AX_MNX);
    const AxAt zp = ax_at(Z, nz, kp, AX_AH, AX_KH, AX_IDLP, AX_MNX);
    const int ly = jp * nz + k - base, lz = j * nz + kp - base;
    const bool y_lds = y_nb && ly >= 0 && ly < OEM_ADJ_TK * OEM_ADJ_TJ;
    const bool z_lds = z_nb && lz >= 0 && lz < OEM_ADJ_TK * OEM_ADJ_TJ;
    // per-thread byte offsets within a plane: the cell, its y and z neighbours, their psi rows
    const uint32_t vo = (uint32_t)(j * nz + k) * EB, vyp = (uint32_t)(jp * nz + k) * EB,
                   vzp = (uint32_t)(j * nz + kp) * EB;
    const uint32_t vpy = (uint32_t)(max(yc.s, 0) * nz + k) * EB;
    const uint32_t vpyp = (uint32_t)(max(yp.s, 0) * nz + k) * EB;
    const uint32_t vpz = (uint32_t)(j * wz + max(zc.s, 0)) * EB;
    const uint32_t vpzp = (uint32_t)(j * wz + max(zp.s, 0)) * EB;
    int buf = 0;

    // the x terms (t3: c1, sgn -1, slot 0; t4: c2, sgn +1, slot 1) of the cell on plane i
    auto qx = [&](int i, F &q3, F &q4) {
        const uint32_t sp = (uint32_t)i * plane * EB;
        const AxAt xc = ax_at_row(X, nx, i, AX_AH, AX_KH, AX_IDLP, AX_MNX);
        const F f1 = mch * -1.f * T::ld(rH, vo, sp + cB);
        const F f2 = mch * 1.f * T::ld(rH, vo, sp + 2u * cB);
        F p0 = T::zero(), p1 = T::zero();
        if (xc.s >= 0) {
            const uint32_t sx = (uint32_t)xc.s * plane * EB;
            p0 = T::ld(rP, vo, off[0] * EB + sx);
            p1 = T::ld(rP, vo, off[1] * EB + sx);
        }
        q3 = xc.s >= 0 ? (xc.ik + xc.a) * f1 + xc.a * p0 : f1;
        q4 = xc.s >= 0 ? (xc.ik + xc.a) * f2 + xc.a * p1 : f2;
    };
    F q3p = T::zero(), q4p = T::zero();
    bool x_nb_bot = false;
    {
        int ipb = i0 - 1;
        if (i0 > 0) {
            x_nb_bot = true;
        } else {
            ipb = X.prv[0];
            x_nb_bot = ipb != 0 && X.nxt[ipb] == 0;
        }
        if (x_nb_bot) qx(ipb, q3p, q4p);
    }
    for (int i = i0; i < i1; ++i) {
        const uint32_t sp = (uint32_t)i * plane * EB;
        const uint32_t sy = (uint32_t)i * (uint32_t)(wy * nz) * EB;
        const uint32_t sz = (uint32_t)i * (uint32_t)(ny * wz) * EB;
        const AxAt xc = ax_at_row(X, nx, i, AX_AH, AX_KH, AX_IDLP, AX_MNX);
        const uint32_t sx = (uint32_t)max(xc.s, 0) * plane * EB;
        const F h0 = T::ld(rH, vo, sp), h1 = T::ld(rH, vo, sp + cB),
                    h2 = T::ld(rH, vo, sp + 2u * cB);
        // f of each term: -ch sgn lH_c
        const F f0p = mch * 1.f * h0, f1m = mch * -1.f * h1, f2p = mch * 1.f * h2;
        const F f0m = mch * -1.f * h0, f1p = mch * 1.f * h1, f2m = mch * -1.f * h2;
        F py0 = T::zero(), py1 = T::zero(), pz0 = T::zero(), pz1 = T::zero(), px0 = T::zero(), px1 = T::zero();
        if (yc.s >= 0) { py0 = T::ld(rP, vpy, off[2] * EB + sy); py1 = T::ld(rP, vpy, off[3] * EB + sy); }
        if (zc.s >= 0) { pz0 = T::ld(rP, vpz, off[4] * EB + sz); pz1 = T::ld(rP, vpz, off[5] * EB + sz); }
        if (xc.s >= 0) { px0 = T::ld(rP, vo, off[0] * EB + sx); px1 = T::ld(rP, vo, off[1] * EB + sx); }
        auto qh = [](const AxAt &t, F f, F psi) {
            return t.s >= 0 ? (t.ik + t.a) * f + t.a * psi : f;
        };
        // every global read of the plane goes out before the barrier (the cell's class and
        // lambda_E for the in-place update, and the neighbours' terms recomputed from their
        // cells where the neighbour's thread is not in the block), so nothing after it waits
        // on a load
        float ax_ = 0.f, ay_ = 0.f, az_ = 0.f;
        F el0 = T::zero(), el1 = T::zero(), el2 = T::zero();
        if (act) {
            cas(vo, sp, (uint32_t)i * plane + (uint32_t)(j * nz + k), ax_, ay_, az_);
            el0 = T::ld(rE, vo, sp);
            el1 = T::ld(rE, vo, sp + cB);
            el2 = T::ld(rE, vo, sp + 2u * cB);
        }
        F q0n = T::zero(), q5n = T::zero(), q1n = T::zero(), q2n = T::zero();
        if (!y_lds && y_nb) {
            F a0 = T::zero(), a1 = T::zero();
            if (yp.s >= 0) { a0 = T::ld(rP, vpyp, off[2] * EB + sy); a1 = T::ld(rP, vpyp, off[3] * EB + sy); }
            q0n = qh(yp, mch * 1.f * T::ld(rH, vyp, sp), a0);
            q5n = qh(yp, mch * -1.f * T::ld(rH, vyp, sp + 2u * cB), a1);
        }
        if (!z_lds && z_nb) {
            F a0 = T::zero(), a1 = T::zero();
            if (zp.s >= 0) { a0 = T::ld(rP, vpzp, off[4] * EB + sz); a1 = T::ld(rP, vpzp, off[5] * EB + sz); }
            q1n = qh(zp, mch * -1.f * T::ld(rH, vzp, sp), a0);
            q2n = qh(zp, mch * 1.f * T::ld(rH, vzp, sp + cB), a1);
        }
        const F q0 = qh(yc, f0p, py0);      // t0: c0 sgn +1, y, slot 2
        const F q5 = qh(yc, f2m, py1);      // t5: c2 sgn -1, y, slot 3
        const F q1 = qh(zc, f0m, pz0);      // t1: c0 sgn -1, z, slot 4
        const F q2 = qh(zc, f1p, pz1);      // t2: c1 sgn +1, z, slot 5
        const F q3 = qh(xc, f1m, px0);      // t3: c1 sgn -1, x, slot 0
        const F q4 = qh(xc, f2p, px1);      // t4: c2 sgn +1, x, slot 1
        s_q[buf][0][tid] = q0;
        s_q[buf][1][tid] = q5;
        s_q[buf][2][tid] = q1;
        s_q[buf][3][tid] = q2;
        __syncthreads();
        if (y_lds) {
            q0n = s_q[buf][0][ly];
            q5n = s_q[buf][1][ly];
        }
        if (z_lds) {
            q1n = s_q[buf][2][lz];
            q2n = s_q[buf][3][lz];
        }
        const bool x_hi = cidx_at(X.nxt, i) == i;
        const bool x_nb = i > i0 || x_nb_bot;
        const int ixp = i > 0 ? i - 1 : cidx_at(X.prv, 0);
        const float mxp = ctab_row(X, ixp)[AX_MNX], mxpi = ctab_row(X, ixp)[AX_MNXI],
                    ixdp = ctab_row(X, ixp)[AX_IDLP];
        // conj(mask), as in the E transpose
        auto term = [](F q, const AxAt &c, bool hi, bool nb, F qn, float mn, float mni,
                       float idn) {
            F v = -q * c.idl;
            if (hi) v += T::mmul(q, c.m, -c.mi) * c.idl;
            if (nb) v += T::mmul(qn, mn, -mni) * idn;
            return v;
        };
        const F v0 = term(q0, yc, y_hi, y_nb, q0n, yp.m, yp.mi, yp.idl);
        const F v1 = term(q1, zc, z_hi, z_nb, q1n, zp.m, zp.mi, zp.idl);
        const F v2 = term(q2, zc, z_hi, z_nb, q2n, zp.m, zp.mi, zp.idl);
        const F v3 = term(q3, xc, x_hi, x_nb, q3p, mxp, mxpi, ixdp);
        const F v4 = term(q4, xc, x_hi, x_nb, q4p, mxp, mxpi, ixdp);
        const F v5 = term(q5, yc, y_hi, y_nb, q5n, yp.m, yp.mi, yp.idl);
        buf ^= 1;
        q3p = q3;
        q4p = q4;
        if (!act) continue;
        F acc0 = T::zero(), acc1 = T::zero(), acc2 = T::zero();
        acc2 += v0; acc1 += v1; acc0 += v2; acc2 += v3; acc1 += v4; acc0 += v5;
        // lambda_E <- ca pec lambda_E (the E update's cell-local part, transposed) + the sum,
        // in place: no thread reads another cell's lambda_E here
        const F e0 = ax_ * (yc.pec * zc.pec) * el0;
        const F e1 = ay_ * (zc.pec * xc.pec) * el1;
        const F e2 = az_ * (xc.pec * yc.pec) * el2;
        T::st(rE, e0 + acc0, vo, sp);
        T::st(rE, e1 + acc1, vo, sp + cB);
        T::st(rE, e2 + acc2, vo, sp + 2u * cB);
        // the cell-local part, into the other buffer: lph <- b (f + lph) per slab term
        if (xc.s >= 0) {
            T::st(rPn, xc.b * (f1m + px0), vo, off[0] * EB + sx);
            T::st(rPn, xc.b * (f2p + px1), vo, off[1] * EB + sx);
        }
        if (yc.s >= 0) {
            T::st(rPn, yc.b * (f0p + py0), vpy, off[2] * EB + sy);
            T::st(rPn, yc.b * (f2m + py1), vpy, off[3] * EB + sy);
        }
        if (zc.s >= 0) {
            T::st(rPn, zc.b * (f0m + pz0), vpz, off[4] * EB + sz);
            T::st(rPn, zc.b * (f1p + pz1), vpz, off[5] * EB + sz);
        }
    }
}

// Objective: Fld[comp][cell] += z * tab[n] for one point list (scale 1; its real part on real
// fields).
template <typename F>
__global__ void __launch_bounds__(256) k_adj_obj(F *Fld, PointsDev pd, MatDev mat, size_t ncell, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pd.n_groups) return;
    point_group<F>(Fld, pd, mat, 1.0f, ncell, n, t);
}

// Every objective point list of a step in one launch, when no two lists touch the same
// (field, component, cell): thread t applies group t - ofs[l] of list l, as k_adj_obj would.
#define OEM_ADJ_OBJ_MAX 8
struct ObjDev {
    int n;
    int ofs[OEM_ADJ_OBJ_MAX + 1];
    int mag[OEM_ADJ_OBJ_MAX];
    PointsDev p[OEM_ADJ_OBJ_MAX];
};
template <typename F>
__global__ void __launch_bounds__(256) k_adj_obj_all(F *lE, F *lH, ObjDev o, MatDev mat,
                                                     size_t ncell, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= o.ofs[o.n]) return;
    int l = 0;
    while (t >= o.ofs[l + 1]) ++l;
    point_group<F>(o.mag[l] ? lH : lE, o.p[l], mat, 1.0f, ncell, n, t - o.ofs[l]);
}

// eps gradient: g[q] += lE . (dca e0 + dcb (e1 - ca e0) / cb); on complex fields
// Re(conj(lE) (...)), the real and imaginary parts each as on real fields
__device__ __forceinline__ double grad_term(float l, float dca, float dcb, float ca, float cb,
                                            float e0, float e1) {
    const float s = (e1 - ca * e0) / cb;
    return (double)l * ((double)dca * e0 + (double)dcb * s);
}
template <typename F>
__global__ void __launch_bounds__(256)
k_adj_grad(const F *lE, const int *comp, const int *cell, const float *dca, const float *dcb,
           const float *ca, const float *cb, const F *e0, const F *e1, int n,
           size_t ncell, double *g)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const F l = lE[(size_t)comp[q] * ncell + cell[q]];
    if constexpr (sizeof(F) == 4) {
        g[q] += grad_term(l, dca[q], dcb[q], ca[q], cb[q], e0[q], e1[q]);
    } else {
        g[q] += grad_term(l.r, dca[q], dcb[q], ca[q], cb[q], e0[q].r, e1[q].r) +
                grad_term(l.i, dca[q], dcb[q], ca[q], cb[q], e0[q].i, e1[q].i);
    }
}

// rec[q] = E[comp[q]][cell[q]]
template <typename F>
__global__ void __launch_bounds__(256)
k_adj_gather(const F *E, const int *comp, const int *cell, int n, size_t ncell, F *rec)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) rec[q] = E[(size_t)comp[q] * ncell + cell[q]];
}

// Pole entry e of either ADE layout: its component, cell and pole range [t0, t1), and pole t's
// coefficients as (A-1 re, A-1 im, B re, B im).
__device__ __forceinline__ void adj_entry(const DispDev &d, int e, uint32_t &c, uint32_t &id,
                                          int &t0, int &t1) {
    if (d.np) {
        const uint2 mt = d.meta[e];
        c = mt.x >> 30;
        id = mt.x & 0x3fffffffu;
        t0 = e * d.np;
        t1 = t0 + d.np;
    } else {
        c = (uint32_t)d.comp[e];
        id = (uint32_t)d.cell[e];
        t0 = d.ofs[e];
        t1 = d.ofs[e + 1];
    }
}
__device__ __forceinline__ float4 adj_coef(const DispDev &d, int t) {
    if (d.np) return d.coef[d.c16[t]];
    const int ci = d.cidx[t];
    return make_float4(d.am1_re[ci], d.am1_im[ci], d.b_re[ci], d.b_im[ci]);
}

// wnu[e] = -cw cb pec of entry e's E component at its cell: lambda_hist = wnu lambda_E
__global__ void __launch_bounds__(256)
k_adj_disp_setup(DispDev d, MatDev mat, GridDev g, float cw, size_t ncell, float *wnu)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    uint32_t c, id;
    int t0, t1;
    adj_entry(d, e, c, id, t0, t1);
    const int pl = g.ny * g.nz;
    const int i = (int)id / pl, r = (int)id - i * pl, j = r / g.nz, k = r - j * g.nz;
    const float cb = mat.mode == 0 ? mat.lut[8 * (size_t)mat.cls[id] + 2 * c + 1]
                                   : mat.cb[(size_t)c * ncell + id];
    const float px = axv(g.ax[0], i, AX_PEC), py = axv(g.ax[1], j, AX_PEC),
                pz = axv(g.ax[2], k, AX_PEC);
    const float pec = c == 0 ? py * pz : c == 1 ? pz * px : px * py;
    wnu[e] = -cw * cb * pec;
}

// The ADE transposes of one backward step, one thread per pole entry: lambda_E^{n+1} gains the
// pending term of lambda_E (from the step after) and the post update's term; then the pre
// update's transpose moves lambda_P and leaves lambda_E^n's term pending.  On complex fields
// the pole state driven by Im E (its adjoint lQ) answers to the imaginary parts, as lP to the
// real ones.
template <typename F>
__global__ void __launch_bounds__(256)
k_adj_disp(F *lE, DispDev d, float2 *lP, float2 *lQ, F *pend, const float *wnu, size_t ncell)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_entry) return;
    uint32_t c, id;
    int t0, t1;
    adj_entry(d, e, c, id, t0, t1);
    F *le = lE + (size_t)c * ncell + id;
    if constexpr (sizeof(F) == 4) {
        float l = *le + pend[e];
        for (int t = t0; t < t1; ++t) {
            const float4 cf = adj_coef(d, t);
            const float2 p = lP[t];
            l += cf.z * p.x + cf.w * p.y;                 // Re(conj(B) lambda_P)
        }
        *le = l;
        const float lh2 = 2.f * wnu[e] * l;               // 2 lambda_hist, on the real part
        float r = 0.f;
        for (int t = t0; t < t1; ++t) {
            const float4 cf = adj_coef(d, t);
            const float2 p = lP[t];
            const float dr = p.x + lh2, di = p.y;         // lambda_dP
            // lambda_P + conj(A-1) lambda_dP
            lP[t] = make_float2(p.x + cf.x * dr + cf.y * di, p.y + cf.x * di - cf.y * dr);
            r += cf.z * dr + cf.w * di;                   // Re(conj(B) lambda_dP)
        }
        pend[e] = r;
    } else {
        cf2 l = *le + pend[e];
        for (int t = t0; t < t1; ++t) {
            const float4 cf = adj_coef(d, t);
            const float2 p = lP[t], q = lQ[t];
            l.r += cf.z * p.x + cf.w * p.y;
            l.i += cf.z * q.x + cf.w * q.y;
        }
        *le = l;
        const float w2 = 2.f * wnu[e];
        const float lhr = w2 * l.r, lhi = w2 * l.i;       // 2 lambda_hist: P gets Re, Q gets Im
        cf2 r = {0.f, 0.f};
        for (int t = t0; t < t1; ++t) {
            const float4 cf = adj_coef(d, t);
            const float2 p = lP[t], q = lQ[t];
            const float dr = p.x + lhr, di = p.y, er = q.x + lhi, ei = q.y;
            lP[t] = make_float2(p.x + cf.x * dr + cf.y * di, p.y + cf.x * di - cf.y * dr);
            lQ[t] = make_float2(q.x + cf.x * er + cf.y * ei, q.y + cf.x * ei - cf.y * er);
            r.r += cf.z * dr + cf.w * di;
            r.i += cf.z * er + cf.w * ei;
        }
        pend[e] = r;
    }
}

}  // namespace

struct oem_adjoint_state {
    // checkpoints
    int n_slots = 0;
    float *ckpt = nullptr;           // n_slots * state_floats (the state of one step)
    std::vector<int64_t> ckpt_step;
    // adjoint fields, of the field type (two floats per value on complex fields)
    float *lE = nullptr, *lH = nullptr, *lpe = nullptr, *lph = nullptr;
    float *lpe2 = nullptr, *lph2 = nullptr;   // the other halves of the ping-pong
    // design
    int n_design = 0, max_seg = 0;
    int *d_comp = nullptr, *d_cell = nullptr;
    float *dca = nullptr, *dcb = nullptr, *ca = nullptr, *cb = nullptr;
    double *grad = nullptr;
    float *rec = nullptr;            // (max_seg + 1) * n_design values of the field type
    int64_t rec_s0 = -1;
    int rec_n = 0;
    // objective point lists: E side, H side; the (field, component, cell) keys they touch, and
    // whether every key is touched by one list only (then one launch applies them all)
    std::vector<Points> obj;
    std::unordered_set<uint64_t> obj_keys;
    bool obj_disjoint = true;
    // ADE poles: lambda_P per pole (lambda_Q of the Im E driven state on complex fields), the
    // pending lambda_E^n term (field type) and -cw cb pec per entry
    float2 *lP = nullptr, *lQ = nullptr;
    float *pend = nullptr, *wnu = nullptr;
};

namespace oemi {
void adjoint_state_free(oem_adjoint_state *a) { delete a; }
}

// x planes per block of the transposed sweeps
static int adj_chunk() {
    const char *e = std::getenv("OEM_ADJ_CHUNK");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : OEM_ADJ_CHUNK;
}

// rec_row[q] = E[comp[q]][cell[q]] of the current state, in the field type
static void gather_design(oem_ctx *ctx, oem_adjoint_state *a, size_t row) {
    const dim3 b(256), g((a->n_design + 255) / 256);
    const size_t o = row * (size_t)a->n_design;
    if (ctx->cplx)
        hipLaunchKernelGGL((k_adj_gather<cf2>), g, b, 0, ctx->stream,
                           reinterpret_cast<const cf2 *>(ctx->cur_E()), a->d_comp, a->d_cell,
                           a->n_design, ctx->ncell, reinterpret_cast<cf2 *>(a->rec) + o);
    else
        hipLaunchKernelGGL((k_adj_gather<float>), g, b, 0, ctx->stream, ctx->cur_E(), a->d_comp,
                           a->d_cell, a->n_design, ctx->ncell, a->rec + o);
}

static int adj_check(oem_ctx *ctx) {
    if (!ctx->prepared) return fail(OEM_ESTATE, "oem_prepare not called");
    if (ctx->pool) return fail(OEM_ENOTSUP, "pooled contexts have no adjoint");
    if ((ctx->dims.nx + adj_chunk() - 1) / adj_chunk() > 65535)
        return fail(OEM_ENOTSUP, "the adjoint launch covers nx <= %d", 65535 * adj_chunk());
    // the transposed sweeps address the three components of an adjoint field, and the psi
    // adjoints of one side, through buffer resources with 32-bit byte offsets
    const uint64_t eb = 4ull * ctx->esz;
    if (3ull * eb * ctx->ncell >= (1ull << 32) || eb * ctx->psi.total >= (1ull << 32))
        return fail(OEM_ENOTSUP, "the adjoint covers up to 2^32 / %llu cells",
                    (unsigned long long)(3ull * eb));
    if (!ctx->adj) ctx->adj = new oem_adjoint_state();
    return OEM_OK;
}

// a checkpoint: E, H, both psi sets and the ADE pole state ((re, im) per pole; between runs it
// holds P^n, the last step's post not fused with a next pre)
static size_t pole_floats(oem_ctx *c) {
    return c->has_disp ? 2 * (size_t)c->disp_n_pole * (c->cplx ? 2 : 1) : 0;
}
static size_t state_floats(oem_ctx *c) {
    return (6 * c->ncell + 2 * (size_t)c->psi.total) * c->esz + pole_floats(c);
}

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
    const size_t np = (size_t)c->disp_n_pole, es = c->esz;
    const bool compact = c->has_disp && c->disp.np, pd = c->has_disp, qd = pd && c->cplx;
    // the pole state: (re, im) pairs (compact) or re and im arrays (CSR), driven by Re E and,
    // on complex fields, by Im E
    float *parts[8] = {E, H, static_cast<float *>(ps.h), static_cast<float *>(ps.e),
                       compact ? reinterpret_cast<float *>(c->disp.sp) : c->disp.p_re,
                       compact ? nullptr : c->disp.p_im,
                       compact ? reinterpret_cast<float *>(c->disp.sq) : c->disp.q_re,
                       compact ? nullptr : c->disp.q_im};
    const size_t n[8] = {3 * nc * es, 3 * nc * es, pt * es, pt * es,
                         pd ? (compact ? 2 * np : np) : 0, pd && !compact ? np : 0,
                         qd ? (compact ? 2 * np : np) : 0, qd && !compact ? np : 0};
    size_t o = 0;
    for (int q = 0; q < 8; ++q) {
        if (n[q] == 0) continue;
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
    if (ctx->pre_ready) return fail(OEM_ESTATE, "checkpoint between runs only");
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
    ctx->pre_ready = false;
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
    const size_t nc = ctx->ncell, es = ctx->esz;
    TRY(ctx->alloc(&a->lE, 3 * nc * es, true));
    TRY(ctx->alloc(&a->lH, 3 * nc * es, true));
    TRY(ctx->alloc(&a->lpe, (size_t)ctx->psi.total * es, true));
    TRY(ctx->alloc(&a->lph, (size_t)ctx->psi.total * es, true));
    TRY(ctx->alloc(&a->lpe2, (size_t)ctx->psi.total * es, true));
    TRY(ctx->alloc(&a->lph2, (size_t)ctx->psi.total * es, true));
    TRY(ctx->upload(&a->d_comp, (const int *)comp, (size_t)n_design));
    TRY(ctx->upload(&a->d_cell, (const int *)cell, (size_t)n_design));
    TRY(ctx->upload(&a->dca, dca, (size_t)n_design));
    TRY(ctx->upload(&a->dcb, dcb, (size_t)n_design));
    TRY(ctx->upload(&a->ca, ca, (size_t)n_design));
    TRY(ctx->upload(&a->cb, cb, (size_t)n_design));
    TRY(ctx->alloc(&a->grad, (size_t)n_design, true));
    TRY(ctx->alloc(&a->rec, (size_t)(max_seg + 1) * n_design * es, false));
    if (ctx->has_disp) {
        const int ne = ctx->disp.n_entry;
        TRY(ctx->alloc(&a->lP, std::max<size_t>((size_t)ctx->disp_n_pole, 1), true));
        if (ctx->cplx) TRY(ctx->alloc(&a->lQ, std::max<size_t>((size_t)ctx->disp_n_pole, 1), true));
        TRY(ctx->alloc(&a->pend, (size_t)ne * es, true));
        TRY(ctx->alloc(&a->wnu, (size_t)ne, true));
        hipLaunchKernelGGL(k_adj_disp_setup, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream,
                           ctx->disp, ctx->mat, ctx->g, ctx->dims.cw, ctx->ncell, a->wnu);
        HIP_OK(hipGetLastError());
    }
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
    oem_adjoint_state *a = ctx->adj;
    std::unordered_set<uint64_t> mine;
    for (int q = 0; q < n_pts; ++q) {
        const uint64_t key = ((uint64_t)(magnetic ? 1 : 0) << 63) | ((uint64_t)comp[q] << 40) |
                             (uint64_t)cell[q];
        mine.insert(key);
    }
    for (uint64_t key : mine)
        if (!a->obj_keys.insert(key).second) a->obj_disjoint = false;
    a->obj.push_back(P);
    return OEM_OK;
}

extern "C" int oem_adjoint_record(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (!a->rec) return fail(OEM_ESTATE, "oem_adjoint_setup not called");
    if (n_steps < 1 || n_steps > a->max_seg) return fail(OEM_EINVAL, "segment of %lld steps", (long long)n_steps);
    if (ctx->step + n_steps > ctx->dims.n_steps) return fail(OEM_EINVAL, "beyond the step tables");
    ctx->mons_off = true;
    a->rec_s0 = ctx->step;
    gather_design(ctx, a, 0);
    int rc = OEM_OK;
    // the last step's ADE post is not fused with a next pre: the pole state ends as P^n, as a
    // run leaves it
    for (int64_t s = 0; s < n_steps && !rc; ++s) {
        rc = oemi::launch_one_step(ctx, s == n_steps - 1);
        if (!rc) gather_design(ctx, a, (size_t)(s + 1));
    }
    ctx->mons_off = false;
    TRY(rc);
    HIP_OK(hipGetLastError());
    a->rec_n = (int)n_steps;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}

extern "C" int oem_adjoint_forward(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (!a->rec) return fail(OEM_ESTATE, "oem_adjoint_setup not called");
    if (n_steps < 1 || n_steps > a->max_seg) return fail(OEM_EINVAL, "segment of %lld steps", (long long)n_steps);
    if (ctx->step + n_steps > ctx->dims.n_steps) return fail(OEM_EINVAL, "beyond the step tables");
    // the cached step graph is captured with or without the recording: drop it on both sides
    if (ctx->graph) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    const int64_t s0 = ctx->step;
    gather_design(ctx, a, 0);
    ctx->rec_hook.comp = a->d_comp;
    ctx->rec_hook.cell = a->d_cell;
    ctx->rec_hook.n = a->n_design;
    ctx->rec_hook.rec = a->rec;
    ctx->rec_hook.base = s0;
    ctx->rec_hook.on = true;
    const int rc = oem_run(ctx, n_steps);
    ctx->rec_hook.on = false;
    if (ctx->graph) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    TRY(rc);
    a->rec_s0 = s0;
    a->rec_n = (int)n_steps;
    return OEM_OK;
}

// The backward steps over the recorded segment, from its last: the objective points, the ADE
// transposes, the eps gradient, then the E and H transposes.  F: the field type.
template <typename F>
static void back_steps(oem_ctx *ctx, oem_adjoint_state *a, AdjDev &d, const ObjDev &od, bool all,
                       dim3 tm, dim3 tb) {
    F *lE = reinterpret_cast<F *>(a->lE), *lH = reinterpret_cast<F *>(a->lH);
    const F *rec = reinterpret_cast<const F *>(a->rec);
    const dim3 gb(256), gg((a->n_design + 255) / 256);
    for (int s = a->rec_n - 1; s >= 0; --s) {
        const int n = (int)(a->rec_s0 + s);          // forward step n: E^n -> E^{n+1}
        if (all) {
            if (od.ofs[od.n] > 0)
                hipLaunchKernelGGL((k_adj_obj_all<F>), dim3((od.ofs[od.n] + 255) / 256), dim3(256),
                                   0, ctx->stream, lE, lH, od, ctx->mat, ctx->ncell, n);
        } else {
            for (auto &P : a->obj)
                hipLaunchKernelGGL((k_adj_obj<F>), dim3((P.d.n_groups + 255) / 256), dim3(256), 0,
                                   ctx->stream, P.magnetic ? lH : lE, P.d, ctx->mat, ctx->ncell, n);
        }
        if (ctx->has_disp)
            hipLaunchKernelGGL((k_adj_disp<F>), dim3((ctx->disp.n_entry + 255) / 256), dim3(256), 0,
                               ctx->stream, lE, ctx->disp, a->lP, a->lQ,
                               reinterpret_cast<F *>(a->pend), a->wnu, ctx->ncell);
        hipLaunchKernelGGL((k_adj_grad<F>), gg, gb, 0, ctx->stream, lE, a->d_comp, a->d_cell,
                           a->dca, a->dcb, a->ca, a->cb, rec + (size_t)s * a->n_design,
                           rec + (size_t)(s + 1) * a->n_design, a->n_design, ctx->ncell, a->grad);
        // E transpose: reads lE and lpe, adds to lH, writes the new lpe to the other buffer;
        // then the H transpose: reads lH and lph, rewrites lE in place as ca pec lE + its
        // sum, writes the new lph to the other buffer
        d.lE = a->lE; d.lpe = a->lpe; d.lpe_n = a->lpe2;
        hipLaunchKernelGGL((k_adj_e1<F>), tm, tb, 0, ctx->stream, d);
        std::swap(a->lpe, a->lpe2);
        d.lph = a->lph; d.lph_n = a->lph2;
        hipLaunchKernelGGL((k_adj_h1<F>), tm, tb, 0, ctx->stream, d);
        std::swap(a->lph, a->lph2);
    }
}

extern "C" int oem_adjoint_back(oem_ctx *ctx, int64_t n_steps) {
    CHECK_CTX(ctx);
    TRY(adj_check(ctx));
    oem_adjoint_state *a = ctx->adj;
    if (a->rec_s0 < 0 || n_steps != a->rec_n)
        return fail(OEM_ESTATE, "back over exactly the recorded segment (%d steps)", a->rec_n);
    AdjDev d{};
    d.lH = a->lH;
    d.psl = ctx->psi;
    d.g = ctx->g;
    d.mat = ctx->mat;
    d.ch = ctx->dims.ch;
    d.ncell = ctx->ncell;
    d.chunk = adj_chunk();
    const dim3 tb(OEM_ADJ_TK, OEM_ADJ_TJ);
    const int cols = ctx->dims.ny * ctx->dims.nz;
    const dim3 tm((cols + OEM_ADJ_TK * OEM_ADJ_TJ - 1) / (OEM_ADJ_TK * OEM_ADJ_TJ), 1,
                  (ctx->dims.nx + d.chunk - 1) / d.chunk);
    // the objective lists in one launch per step when they touch disjoint cells
    ObjDev od{};
    const bool all = a->obj_disjoint && a->obj.size() <= OEM_ADJ_OBJ_MAX &&
                     !(std::getenv("OEM_ADJ_OBJ_ALL") && std::getenv("OEM_ADJ_OBJ_ALL")[0] == '0');
    if (all) {
        od.n = (int)a->obj.size();
        for (int l = 0; l < od.n; ++l) {
            od.p[l] = a->obj[l].d;
            od.mag[l] = a->obj[l].magnetic ? 1 : 0;
            od.ofs[l + 1] = od.ofs[l] + a->obj[l].d.n_groups;
        }
    }
    if (ctx->cplx)
        back_steps<cf2>(ctx, a, d, od, all, tm, tb);
    else
        back_steps<float>(ctx, a, d, od, all, tm, tb);
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
    HIP_OK(hipMemcpyAsync(out, a->grad, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return OEM_OK;
}
