// Host-side internals shared by the single-scene library (openem_hip.hip) and the serving pool
// (oem_pool.hip): error reporting, the resident-scene context, and the arena hooks a pool
// context allocates through.

#pragma once

#include "../../include/openem_hip.h"
#include "oem_kernels.h"

#include <mutex>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace oemi {

// One message per host thread; the C ABI returns a code and oem_last_error reads this.
inline thread_local std::string g_err;

inline int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct Monitor {
    MonDev d;
    size_t count;         // ncomp*nf*npts
    bool fused = false;   // the fused step writes its snapshots; `patch` lists the points an
    int *patch = nullptr; // E-side point list corrects afterwards, snapped again
    int n_patch = 0;
};

struct TimeMonitor {
    TimeMonDev d;
    size_t count;     // nc*slots*cells elements of the field type
};

struct FluxTimeMonitor {
    FluxTimeDev d;
};

// One grouped point list (dipoles or TF/SF face corrections) and which field it updates.
enum { PT_DIPOLE = 0, PT_TFSF = 1, PT_MODE = 2 };

// Smallest scene (cells) that takes the fused H+E step by default.
#define OEM_FUSE_MIN_CELLS (10u << 20)

struct Points {
    PointsDev d;
    int magnetic;       // 1: applied to H after update_h; 0: to E after update_e
    int kind;           // PT_*: launch order H: dipole, TF/SF, mode; E: TF/SF, mode, dipole
    std::vector<int> cells;   // the group cells (host copy)
};

// Kernel family of a point list (profiling ids of include/openem_hip.h).
inline int points_kid(const Points &p) {
    if (p.kind == PT_MODE) return p.magnetic ? OEM_K_MODE_H : OEM_K_MODE_E;
    if (p.kind == PT_TFSF) return p.magnetic ? OEM_K_TFSF_H : OEM_K_TFSF_E;
    return p.magnetic ? OEM_K_DIPOLE_H : OEM_K_DIPOLE_E;
}

// Position of a point list in the step (0..5): H side dipole, TF/SF, mode; E side TF/SF,
// mode, dipole -- the reference's order (`openem/solver.py:876-968`).
inline int points_slot(const Points &p) {
    static const int h[3] = {0, 1, 2}, e[3] = {5, 3, 4};
    return p.magnetic ? h[p.kind] : e[p.kind];
}

// Arena of a serving pool (oem_pool.hip).
int pool_arena_alloc(oem_pool *pool, size_t bytes, void **out);
void pool_arena_free(oem_pool *pool, void *p);

}  // namespace oemi

namespace oemi {
// Held while a stream is captured into a graph (begin to instantiate) and around every call
// that may synchronise the device (hipFree, hipHostFree, hipStreamDestroy, hipMalloc):
// contexts step on several host threads at once (serving lanes, designs in flight), and a
// device-wide synchronisation issued by one thread while another thread's stream is capturing
// invalidates that capture.  For the same reason no copy here goes through the legacy stream:
// every host copy is hipMemcpyAsync on the context's (or pool's) own stream, then a
// synchronisation of that stream.
std::mutex &device_sync_mutex();
}

struct oem_reduce_state;            // energy / flux reduction tables (oem_reduce.hip)
struct oem_adjoint_state;           // checkpoints, adjoint fields, design gradient (oem_adjoint.hip)
namespace oemi {
void reduce_state_free(oem_reduce_state *r);
void adjoint_state_free(oem_adjoint_state *a);
}

using oemi::fail;
using oemi::Monitor;
using oemi::TimeMonitor;
using oemi::FluxTimeMonitor;
using oemi::Points;
using oemi::PT_DIPOLE;
using oemi::PT_MODE;
using oemi::PT_TFSF;

#define HIP_OK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? OEM_ENOMEM : OEM_EHIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_) return rc_;       \
    } while (0)

struct oem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    oem_dims dims{};
    size_t ncell = 0;
    bool cplx = false;          // Bloch: complex fields, (re, im) interleaved
    size_t esz = 1;             // floats per field element
    float *E = nullptr, *H = nullptr;
    PsiDev psi{};
    // fused H+E step: a second set of fields and psi; after s steps the state is in set s % 2
    bool fused = false;
    float *E2 = nullptr, *H2 = nullptr;
    PsiDev psi2{};
    float *cur_E() const { return (fused && (step & 1)) ? E2 : E; }
    float *cur_H() const { return (fused && (step & 1)) ? H2 : H; }
    GridDev g{};
    GridDev fg{};               // launch geometry of the fused H+E step
    FuseSnap fsnap{};           // the DFT monitors the fused step snaps
    bool fsnap_multi = false;   // some work-group meets more than one of them
    bool axis_set[3] = {false, false, false};
    MatDev mat{};
    bool mat_set = false;
    std::vector<char> class_flagged;   // LUT classes that read the ADE history term
    DispDev disp{};
    bool has_disp = false;
    int64_t disp_n_pole = 0;           // poles over all entries (the pole state's length)
    PlaneSrcDev pw{};
    std::vector<Points> points;
    std::vector<Monitor> mons;
    std::vector<TimeMonitor> tmons;
    std::vector<FluxTimeMonitor> ftmons;
    int *d_step = nullptr;      // [0] steps taken, [1] steps flushed into the DFT accumulators
    int64_t step = 0;
    int pending = 0;
    bool pre_ready = false;     // ADE: hist/P already advanced by the previous fused step
    bool prepared = false;
    int chunk_req = 0;          // x-planes per work-group requested (0: heuristic)
    int form_req = 0;           // step form requested: 0 the library's choice, 1 two sweeps, 2 fused
    bool mons_off = false;      // adjoint recomputation: no DFT snapshots or flushes
    // adjoint forward recording (oem_adjoint_forward): after each step, E at the design cells
    // goes to row (step + 1 - base) of rec
    struct RecHook {
        bool on = false;
        const int *comp = nullptr, *cell = nullptr;
        int n = 0;
        float *rec = nullptr;       // of the field type (two floats per value on complex fields)
        int64_t base = 0;
    } rec_hook;
    struct oem_adjoint_state *adj = nullptr;
    hipGraphExec_t graph = nullptr;
    std::vector<void *> allocs;
    oem_reduce_state *red = nullptr;

    // serving pool membership (oem_pool.hip); pool == nullptr for a standalone context
    oem_pool *pool = nullptr;
    bool attached = false;
    int64_t done_block = -1;    // pool block in which the last step runs (-1: not scheduled)
    int64_t retire_after = -1;  // memory is reclaimed once this pool block has completed

    ~oem_ctx() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (red) oemi::reduce_state_free(red);
        if (adj) oemi::adjoint_state_free(adj);
        std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
        for (void *p : allocs) {
            if (pool) oemi::pool_arena_free(pool, p);
            else (void)hipFree(p);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    int alloc(T **out, size_t n, bool zero) {
        void *p = nullptr;
        if (n == 0) n = 1;
        if (pool) {
            TRY(oemi::pool_arena_alloc(pool, n * sizeof(T), &p));
        } else {
            std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
            HIP_OK(hipMalloc(&p, n * sizeof(T)));
        }
        allocs.push_back(p);
        if (zero) HIP_OK(hipMemsetAsync(p, 0, n * sizeof(T), stream));
        *out = static_cast<T *>(p);
        return OEM_OK;
    }

    template <typename T>
    int upload(T **out, const T *host, size_t n) {
        int rc = alloc(out, n, false);
        if (rc) return rc;
        HIP_OK(hipMemcpyAsync(*out, host, n * sizeof(T), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return OEM_OK;
    }
};

#define CHECK_CTX(c)                                                           \
    do {                                                                       \
        if (!(c)) return fail(OEM_EINVAL, "null context");                     \
        HIP_OK(hipSetDevice((c)->device));                                     \
    } while (0)

// Shared by oem_ctx_create and oem_pool_ctx_create: validate dims and allocate the fields,
// psi slabs and counters of a context (through the pool's arena when ctx->pool is set).
int ctx_init(oem_ctx *c, const oem_dims *d, int32_t device);

// Pool hooks used by the single-scene entry points (oem_pool.hip), and the point-list builder
// and step launcher shared with the adjoint (openem_hip.hip).
namespace oemi {
int build_points(oem_ctx *ctx, int magnetic, int kind, int tfsf, int cplx_amp, int32_t n_pts,
                 const int32_t *comp, const int32_t *cell, const float *coef, const int32_t *col,
                 const float *phase, int32_t ncol, const double *tab, size_t tab_doubles,
                 Points *out);
int launch_one_step(oem_ctx *ctx, bool last);       // one eager step, no flush, no profiling
int pool_retire(oem_ctx *ctx);                  // oem_ctx_destroy of a pool context
int pool_wait_done(oem_ctx *ctx);               // before reading a pool context's results
}
