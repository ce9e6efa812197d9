/*
 * openem_hip.h - C ABI of the MI355X (gfx950) FDTD time-stepper.
 *
 * The reference drives its CUDA kernels from Python through cupy (`openem/device.py:145-170`,
 * "the only place that imports cupy ... changing to ctypes ... changes only this one file",
 * `openem/device.py:3-10`) and builds every device table in `openem/solver.py:403-489`.
 * This header is the replacement boundary: plain pointers and sizes, integer status codes,
 * no exceptions, the caller owns host buffers, the library owns device buffers.
 *
 * Life cycle: oem_ctx_create -> oem_set_axis x3 -> oem_set_materials_* -> optional tables
 * (dispersion, plane sources, dipoles, monitors) -> oem_prepare -> oem_run ... -> oem_read_*
 * -> oem_ctx_destroy.  A context is used from one host thread at a time.
 *
 * Layout: fields are C order (nx, ny, nz) float32, z contiguous, flat cell (i*ny + j)*nz + k,
 * the reference's flat index (`openem/grid.py:92-103`) with no pitch padding.
 */
#ifndef OPENEM_HIP_H
#define OPENEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OEM_ABI_VERSION 1

enum oem_status {
    OEM_OK = 0,
    OEM_EINVAL = 1,   /* bad argument or table mismatch   (reference: ValueError)           */
    OEM_ENOMEM = 2,   /* device allocation failed                                            */
    OEM_EHIP = 3,     /* HIP runtime error                                                   */
    OEM_ESTATE = 4,   /* call out of order (e.g. oem_run before oem_prepare)                 */
    OEM_ENOTSUP = 5   /* unsupported physics  (reference: NotImplementedError)               */
};

/* Kernel families reported by oem_profile_run (index into its output arrays). */
enum oem_kernel_id {
    OEM_K_UPDATE_H = 0, OEM_K_UPDATE_E = 1, OEM_K_DIPOLE_E = 2, OEM_K_DIPOLE_H = 3,
    OEM_K_DISP_PRE = 4, OEM_K_DISP_POST = 5, OEM_K_DFT_SNAP = 6, OEM_K_DFT_FLUSH = 7,
    OEM_K_STEP = 8, OEM_K_TFSF_H = 9, OEM_K_TFSF_E = 10, OEM_K_MODE_H = 11, OEM_K_MODE_E = 12,
    OEM_K_UPDATE_HE = 13, OEM_K_STEP_TAIL = 14, OEM_K_COUNT = 15
};

typedef struct oem_ctx oem_ctx;
typedef struct oem_pool oem_pool;

typedef struct oem_dims {
    int32_t nx, ny, nz;        /* cells (PML layers included)                                 */
    int32_t n_steps;           /* length of every per-step table (>= steps ever run)          */
    int32_t pml_lo[3];         /* CPML slab thickness per axis at the low / high end;          */
    int32_t pml_hi[3];         /*   psi is stored only inside these slabs                     */
    int32_t complex_fields;    /* 1 = Bloch: complex fields, (re, im) float pairs per element */
    int32_t dft_batch;         /* DFT snapshot depth T (steps per fp64 flush), 1..64          */
    float ch;                  /* dt / mu0                                                    */
    float cw;                  /* eps0 / dt                                                   */
} oem_dims;

/* Library version (OEM_ABI_VERSION).  Replaces nothing: handshake for the ctypes binding. */
int oem_version(void);

/* Copy the calling thread's last error message (NUL-terminated, truncated to n). */
int oem_last_error(char *buf, size_t n);

/* Number of visible HIP devices (0 on a machine without a GPU; never an error). */
int oem_device_count(int32_t *n);

/* Allocate fields E, H (zeroed) and the CPML psi slabs on `device`.
 * Replaces device_tables._alloc_fields (`openem/device_tables.py:42-86`). */
int oem_ctx_create(const oem_dims *d, int32_t device, oem_ctx **out);
int oem_ctx_destroy(oem_ctx *ctx);

/* One axis (0=x,1=y,2=z): 1/dl primal and dual, neighbour tables nxt/prv (int32), masks
 * mnx/mpv/pec, CPML (a, b, 1/kappa) at E and H positions, each of length n with the identity
 * (0, 1, 1) outside the PML; mnx_im/mpv_im (may be NULL) are the imaginary parts of the masks,
 * non-zero only at a Bloch wrap (complex contexts only).  Replaces grid.Axis.index_tables
 * (`openem/grid.py:219-304`), coeffs._pml_axis_arrays (`openem/coeffs.py:21-45`),
 * device_tables._axis_device_tables (`openem/device_tables.py:118-165`) and the complex masks
 * of the Bloch path (`openem/kernels/bloch.cu:3-23`). */
int oem_set_axis(oem_ctx *ctx, int32_t axis, const float *idl_p, const float *idl_d,
                 const int32_t *nxt, const int32_t *prv, const float *mnx, const float *mpv,
                 const float *pec, const float *a_e, const float *b_e, const float *k_e,
                 const float *a_h, const float *b_h, const float *k_h, const float *mnx_im,
                 const float *mpv_im);

/* Material classes: one uint16 class per cell and a LUT of n_classes rows of 8 floats
 * (ca_x, cb_x, ca_y, cb_y, ca_z, cb_z, flags as uint32 bits 0..2 = "component is dispersive",
 * pad).  Replaces the (ca, cb) tables and the uint16 LUT of coeffs._e_coeffs / _cacb_lut
 * (`openem/coeffs.py:98-136`, `:174-208`). */
int oem_set_materials_lut(oem_ctx *ctx, const uint16_t *cls, int32_t n_classes, const float *lut);

/* Dense fallback: ca, cb each 3*ncell floats (component-major).  Replaces
 * device_tables._e_coeff_tables (`openem/device_tables.py:168-219`). */
int oem_set_materials_dense(oem_ctx *ctx, const float *ca, const float *cb);

/* Complex-pole ADE in CSR form: entry e -> component comp[e] at flat cell cell[e], poles
 * ofs[e]..ofs[e+1]; per pole an index cidx into the (A-1, B) coefficient LUT; hist_box
 * (x0,x1,y0,y1,z0,z1) bounds all entry cells.  Replaces device_tables._dispersion_tables
 * (`openem/device_tables.py:235-285`) and the dispersion_pre/post launches
 * (`openem/solver.py:920-1024`). */
int oem_set_dispersion(oem_ctx *ctx, int32_t n_entry, const int32_t *comp, const int32_t *cell,
                       const int32_t *ofs, int32_t n_pole, const int32_t *cidx, int32_t n_coef,
                       const float *am1_re, const float *am1_im, const float *b_re,
                       const float *b_im, const int32_t *hist_box);

/* One-face TF/SF plane wave along z (pol 0: Ex/Hy, 1: Ey/Hx), at most 2 per context.
 * ex_inc / hy_inc have n_steps+1 float64 samples.  Replaces sources_setup._plane_source_one
 * (`openem/sources_setup.py:93-119`) and the inject_h / inject_e launches
 * (`openem/kernels/source_dft.cu:40-73`). */
int oem_add_plane_source(oem_ctx *ctx, int32_t pol, int32_t kh, int32_t ks, float coef_h,
                         float coef_e, const double *ex_inc, const double *hy_inc);

/* Point dipoles (electric: E -= cb*coef*amp[src][n]; magnetic: H -= ch*coef*amp[src][n]).
 * amp is n_src rows of n_steps+1 float64 samples, (re, im) pairs when amp_complex (complex
 * contexts inject the complex waveform, real ones its real part).  Points on the same
 * (comp, cell) are applied by one thread in input order (no atomics, deterministic).
 * Replaces sources_setup._dipole_table (`openem/sources_setup.py:419-457`, `:521`) and
 * inject_dipole / inject_dipole_h (`openem/kernels/source_dft.cu:147-206`,
 * `openem/kernels/bloch.cu:226-278`). */
int oem_add_dipoles(oem_ctx *ctx, int32_t magnetic, int32_t n_pts, const int32_t *comp,
                    const int32_t *cell, const float *coef, const int32_t *src_of,
                    int32_t n_src, const double *amp, int32_t amp_complex);

/* TF/SF box face corrections as a point list: F[comp][cell] += scale*coef*(table[n][col]*phase)
 * with scale = dt/mu0 for H points (magnetic=1, applied after update_h with the incident E at
 * time n) and the cell's cb for E points (applied after update_e with the incident H at
 * n+1/2).  table is (n_steps+1) x ncol complex float64 (re, im) from the 1D auxiliary line;
 * phase is the per-point transverse Bloch phase (re, im) float32 (imaginary parts must be 0 on
 * real contexts).  Replaces _tfsf_box / tfsf1d.tables (`openem/sources_setup.py:236`,
 * `openem/tfsf1d.py:327`) and tfsf_corr_h / tfsf_corr_e (`openem/kernels/tfsf.cu:49-83`). */
int oem_add_tfsf_points(oem_ctx *ctx, int32_t magnetic, int32_t n_pts, const int32_t *comp,
                        const int32_t *cell, const float *coef, const int32_t *col,
                        const float *phase, int32_t ncol, const double *table);

/* Mode source points on one plane: H side (magnetic = 1, after the TF/SF H corrections)
 * H[comp][cell] += ch * coef * Re(phase * amp[n]) with ch = dt/mu0 and amp the whole-step
 * amplitude table, E side (magnetic = 0, after the TF/SF E corrections)
 * E[comp][cell] += cb * coef * Re(phase * amp[n]) with the cell's cb and the half-step table;
 * phase (re, im float pairs) is the mode profile value at the point and coef carries the face
 * sign and 1/dl.  amp: n_steps+1 complex (re, im
 * doubles).  Real fields take the real part of a complex profile.  Replaces inject_mode_h /
 * inject_mode_e and _launch_mode_inject (`openem/kernels/source_dft.cu:84-128`,
 * `openem/solver.py:230-268`). */
int oem_add_mode_points(oem_ctx *ctx, int32_t magnetic, int32_t n_pts, const int32_t *comp,
                        const int32_t *cell, const float *coef, const float *phase,
                        const double *amp);

/* Plane DFT monitor (normal `axis`, E plane index km >= 1): accumulates
 * (E_t1, E_t2, (H_t1[km]+H_t1[km-1])/2, (H_t2..)/2) for nf frequencies in float64.
 * phase is n_steps*nf*4 float64: cos/sin at t_e=(n+1)dt then at t_h=(n+1/2)dt, windowed.
 * Replaces monitors_setup._flux_mon_entry (`openem/monitors_setup.py:162-208`) and
 * accumulate_dft / dft_snap_plane / dft_flush_plane (`openem/kernels/source_dft.cu:371-430`). */
int oem_add_flux_plane(oem_ctx *ctx, int32_t axis, int32_t km, int32_t nf, const double *phase,
                       int32_t *mon_id);

/* Box DFT monitor, all six components at native Yee positions.  Replaces
 * monitors_setup._field_mon_entry (`openem/monitors_setup.py:211-256`) and
 * accumulate_dft_box / dft_snap_box / dft_flush_box (`openem/kernels/source_dft.cu:216-271`). */
int oem_add_field_box(oem_ctx *ctx, const int32_t *origin, const int32_t *box, int32_t nf,
                      const double *phase, int32_t *mon_id);

/* Time-domain field monitor: components comps[0..n_comps) (0..5 = Ex..Hz, at most 6) in the
 * index box origin/box, sampled at steps m = step_begin, step_begin + interval, ... (m <
 * step_end, at most n_slots): E^m and (H^{m-1/2} + H^{m+1/2}) / 2.  Buffer layout (n_comps,
 * n_slots, ni, nj, nk), zero-initialised.  Replaces monitors_setup._tmon_entry
 * (`openem/monitors_setup.py:258-286`) and sample_time_box (`openem/kernels/source_dft.cu:272-316`). */
int oem_add_time_monitor(oem_ctx *ctx, int32_t n_comps, const int32_t *comps,
                         const int32_t *origin, const int32_t *box, int32_t step_begin,
                         int32_t step_end, int32_t interval, int32_t n_slots, int32_t *mon_id);

/* The whole time-monitor buffer: count = n_comps * n_slots * cells floats (twice that, (re,
 * im) pairs, on a complex context).  Replaces readout._readout's time_samples
 * (`openem/readout.py:183-210`). */
int oem_read_time_monitor(oem_ctx *ctx, int32_t mon_id, float *out, size_t count);

/* Time-domain flux monitor (real fields): one float64 scalar per slot, sum_p w_p sum dS (E_u
 * Hbar_v - E_v Hbar_u) with the sampling window of a time monitor.  Geometry from the host:
 * per transverse axis (ascending t1, t2) n samples with their edge index e, lower / upper
 * cell l / h, centre -> edge weight w of the lower cell and trapezoid weight tr; per normal
 * plane (1 or 2) the weight plane_w, the E layer ke, the lower H cell kh and the coordinate
 * weights wh_lo / wh_hi of H at kh / kh+1; sign = normal_dir (times -1 for a y normal).
 * Replaces monitors_setup._ftmon_entry (`openem/monitors_setup.py:36-71`) and
 * readout._sample_flux_time (`openem/readout.py:20-73`). */
int oem_add_flux_time_monitor(oem_ctx *ctx, int32_t axis, int32_t n1, const int32_t *e1,
                              const int32_t *l1, const int32_t *h1, const float *w1,
                              const double *tr1, int32_t n2, const int32_t *e2,
                              const int32_t *l2, const int32_t *h2, const float *w2,
                              const double *tr2, int32_t n_planes, const double *plane_w,
                              const int32_t *ke, const int32_t *kh, const float *wh_lo,
                              const float *wh_hi, double sign, int32_t step_begin,
                              int32_t step_end, int32_t interval, int32_t n_slots,
                              int32_t *mon_id);

/* The flux series of a flux-time monitor: count = n_slots doubles. */
int oem_read_flux_time_monitor(oem_ctx *ctx, int32_t mon_id, double *out, size_t count);

/* Launch shape: x-planes each work-group marches (0 = the library's choice).  Before
 * oem_prepare.  No counterpart in the reference (its launch shapes are fixed in the kernels'
 * callers, `openem/solver.py:876-1162`, and its block size is the hand-set BLOCK3 knob,
 * `openem/knobs.py:64`); set by the launch autotuner (openem_amd/autotune.py). */
int oem_set_chunk(oem_ctx *ctx, int32_t chunk);

/* Step form: 0 = the library's choice (the fused H+E march from 10 M cells, else the two
 * sweeps; OEM_FUSE=0/1 forces one), 1 = the two sweeps, 2 = the fused march (OEM_ENOTSUP when a
 * field set exceeds its 32-bit offsets or the context is pooled).  Before oem_prepare.  No
 * counterpart in the reference (its H+E fusion is a global knob, `openem/knobs.py:61`); set by
 * the launch autotuner (openem_amd/autotune.py) together with oem_set_chunk. */
int oem_set_step_form(oem_ctx *ctx, int32_t form);

/* The launch plan oem_prepare chose: form 1 (two sweeps) or 2 (fused) and its x chunk. */
int oem_launch_plan(oem_ctx *ctx, int32_t *form, int32_t *chunk);

/* Freeze the tables, build the launch plan.  Must precede oem_run. */
int oem_prepare(oem_ctx *ctx);

/* Advance n steps (graph-captured blocks of dft_batch steps where possible) and flush the
 * DFT accumulators.  Replaces solver._time_loop (`openem/solver.py:1165-1213`). */
int oem_run(oem_ctx *ctx, int64_t n_steps);

/* As oem_run but launched kernel by kernel with HIP events on the context stream around
 * every launch; ms[k] receives the summed device time of kernel family k and calls[k] the
 * launch count (arrays of OEM_K_COUNT).  Replaces device.Profiler (`openem/device.py:29-139`). */
int oem_profile_run(oem_ctx *ctx, int64_t n_steps, double *ms, int64_t *calls);

/* Wall time of oem_run(ctx, n_steps) measured with HIP events on the context stream. */
int oem_timed_run(oem_ctx *ctx, int64_t n_steps, double *ms);

/* Steps completed so far. */
int oem_step_index(oem_ctx *ctx, int64_t *n);

/* Monitor accumulators: count = ncomp*nf*npoints doubles each, layout (comp, f, points) with
 * points in C order of the plane (n1, n2) or the box (ni, nj, nk). */
int oem_read_monitor(oem_ctx *ctx, int32_t mon_id, double *re, double *im, size_t count);

/* Field component 0..5 (Ex Ey Ez Hx Hy Hz) to/from host: ncell floats, or 2*ncell (re, im)
 * pairs on a complex context. */
int oem_read_field(oem_ctx *ctx, int32_t comp, float *out, size_t count);
int oem_write_field(oem_ctx *ctx, int32_t comp, const float *in, size_t count);

/* ---- Float64 reductions: shutoff energies and plane flux -----------------------------
 * Tables for the volume-weighted energy sums: per axis the primal and dual cell lengths (n
 * doubles each) and eps per E component, either per material class (eps_per_class = 1:
 * n_classes rows of 4 floats, eps_x eps_y eps_z pad, LUT materials only) or dense (3 * ncell
 * floats).  Replaces device_tables._dual_volumes_eps (`openem/device_tables.py:320-343`). */
int oem_set_energy_tables(oem_ctx *ctx, const double *dl_x, const double *dl_y,
                          const double *dl_z, const double *dd_x, const double *dd_y,
                          const double *dd_z, const float *eps, int32_t eps_per_class);

/* out[0] = sum V|E|^2, out[1] = sum V eps |E|^2, out[2] = sum V |H|^2 over every cell and
 * component (V = the E component's Yee dual volume as float32, each term formed in float32,
 * accumulated in float64, deterministic order).  Replaces solver._sum2 inside _shutoff_check
 * (`openem/solver.py:107-122`, `:162-177`). */
int oem_energy(oem_ctx *ctx, double *out);

/* Face weights (npts doubles each, the monitor's (n1, n2) layout) for oem_monitor_flux. */
int oem_set_flux_weights(oem_ctx *ctx, int32_t mon_id, const double *w1, const double *w2);

/* Flux terms of a plane monitor from its float64 accumulators, reduced on the device:
 * out[2f] = sum w1 Re(E_t1 conj H_t2), out[2f+1] = sum w2 Re(E_t2 conj H_t1); the flux is
 * (out[2f] - out[2f+1]) / 2.  count = 2 * nf.  Replaces the host reduction of flux.plane_flux
 * (`openem/flux.py:75-110`) on the shutoff and readout paths. */
int oem_monitor_flux(oem_ctx *ctx, int32_t mon_id, double *out, size_t count);

/* oem_monitor_flux without any device allocation, for callers that cannot allocate (serving
 * loops that must not stall the device with a first-call hipMalloc, harnesses that forbid
 * allocation).  No counterpart in the reference.  Same arguments, checks and bits as
 * oem_monitor_flux; the partial sums live in one file-scope device array per GPU (1.04 MB,
 * held until the process ends), so monitors of at most 1024 frequencies are taken (more:
 * OEM_EINVAL).  Calls on one GPU take turns on that array (a host mutex per GPU). */
int oem_monitor_flux_noalloc(oem_ctx *ctx, int32_t mon_id, double *out, size_t count);

/* Overlap weights of a plane monitor: n_sets (1..16) sets of 4 * nf * npts complex values
 * (re, im doubles), layout (set, component, f, point) over the monitor's (E_t1, E_t2, H_t1,
 * H_t2) accumulators. */
int oem_set_overlap_weights(oem_ctx *ctx, int32_t mon_id, int32_t n_sets, const double *g);

/* out[(set * nf + f) * 2 + {0, 1}] = sum over points and components of accumulator * weight
 * (complex, float64, deterministic).  With the conjugated mode profiles as weights these are
 * the mode overlaps.  Replaces the mode-overlap readout of nb/backend._mode_data
 * (`openem/nb/backend.py:1422`), done there on the host. */
int oem_monitor_overlap(oem_ctx *ctx, int32_t mon_id, double *out, size_t count);

/* ---- Discrete adjoint (real and complex Bloch fields, ADE poles included) ------------
 * The reference hands differentiation to tidy3d's adjoint (`openem/nb/autograd_hook.py:1-28`,
 * gradients assembled in `autograd_hook._vjp`); here the exact transpose of the time step runs
 * on the device.  Use: checkpoint the forward run every K steps, then per segment from the last:
 * oem_checkpoint_load, oem_adjoint_record (recompute it, E recorded at the design cells),
 * oem_adjoint_back (the transposed steps; the objective's dJ/dfield and the eps gradient are
 * applied on the way). */
int oem_checkpoint_alloc(oem_ctx *ctx, int32_t n_slots);
/* E, H, psi, the ADE pole state and the step index (between runs: after an oem_run call). */
int oem_checkpoint_save(oem_ctx *ctx, int32_t slot);
int oem_checkpoint_load(oem_ctx *ctx, int32_t slot);
/* Design entries (component 0..2, flat cell) with d ca/d eps, d cb/d eps, ca, cb (float32 each)
 * and the longest segment to be recorded. */
int oem_adjoint_setup(oem_ctx *ctx, int32_t n_design, const int32_t *comp, const int32_t *cell,
                      const float *dca, const float *dcb, const float *ca, const float *cb,
                      int32_t max_seg);
/* The objective's gradient with respect to the field samples of forward step n:
 * lambda[comp][cell] += z * tab[n] (z: (re, im) floats per point, tab: n_steps complex
 * doubles; the real part on real fields, dJ/d Re + i dJ/d Im on complex fields); magnetic = 1
 * for H^{n+1/2}, 0 for E^{n+1}. */
int oem_adjoint_objective(oem_ctx *ctx, int32_t magnetic, int32_t n_pts, const int32_t *comp,
                          const int32_t *cell, const float *z, const double *tab);
/* Recompute n_steps forward from the loaded checkpoint (no DFT), recording E at the design. */
int oem_adjoint_record(oem_ctx *ctx, int64_t n_steps);
/* The forward run itself, n_steps with the monitors (as oem_run), recording E at the design
 * after every step, so no checkpoint or recomputation is needed when the record ((n_steps + 1)
 * x n_design values of the field type, allocated by oem_adjoint_setup's max_seg) fits in
 * memory.  The
 * reference's counterpart is the same: tidy3d's forward run feeds its adjoint
 * (`openem/nb/autograd_hook.py:1-28`). */
int oem_adjoint_forward(oem_ctx *ctx, int64_t n_steps);
/* The transposed steps over the recorded segment, last step first. */
int oem_adjoint_back(oem_ctx *ctx, int64_t n_steps);
/* dJ/d eps of every design entry (float64). */
int oem_adjoint_gradient(oem_ctx *ctx, double *out, size_t n);

/* ---- Serving pool: many resident scenes advanced by the same launches ------------------
 * No counterpart in the reference, which runs one scene per openem.solver.run call
 * (`openem/solver.py:403-511`) with its own allocations.  A pool owns one device arena of
 * arena_bytes (the admission budget) and a hipGraph of dft_batch steps whose kernels, on
 * fixed over-sized grids (blocks past the work exit), walk the tiles, point lists, ADE entries
 * and monitors of every attached scene; each
 * scene keeps its own device step counter, so scenes at different steps share every launch and
 * one that reaches its n_steps simply stops.  Contexts made by oem_pool_ctx_create are built
 * and uploaded with the single-scene calls above (from any host thread), then attached. */
int oem_pool_create(int32_t device, uint64_t arena_bytes, int32_t dft_batch, oem_pool **out);

/* Fails with OEM_ESTATE while contexts of the pool are alive. */
int oem_pool_destroy(oem_pool *pool);

/* As oem_ctx_create, with every device buffer carved out of the pool's arena; OEM_ENOMEM when
 * the arena cannot hold it (nothing stays allocated). */
int oem_pool_ctx_create(oem_pool *pool, const oem_dims *d, oem_ctx **out);

/* Join the batch from the next block on.  The context must be prepared, at step 0, of real
 * fields, with dft_batch equal to the pool's and at most one point list per kind. */
int oem_pool_attach(oem_pool *pool, oem_ctx *ctx);

/* Enqueue n_blocks blocks of dft_batch steps for every attached scene that has steps left;
 * returns without waiting. */
int oem_pool_launch(oem_pool *pool, int32_t n_blocks);

/* *done = 1 once every step of an attached context has run on the device (wait = 1 blocks
 * until then; OEM_ESTATE if the launches enqueued so far do not finish it). */
int oem_pool_ctx_done(oem_pool *pool, oem_ctx *ctx, int32_t wait, int32_t *done);

/* Wait until at most max_pending enqueued blocks are unfinished (0: all finished). */
int oem_pool_wait(oem_pool *pool, int32_t max_pending);

/* Arena bytes in use, arena capacity, attached scenes with steps left, blocks launched. */
int oem_pool_stats(oem_pool *pool, uint64_t *used, uint64_t *capacity, int32_t *n_running,
                   int64_t *blocks);

#ifdef __cplusplus
}
#endif
#endif /* OPENEM_HIP_H */
