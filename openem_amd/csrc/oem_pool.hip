// Serving pool (include/openem_hip.h, oem_pool_*): many resident scenes advanced together.
//
// Device memory.  A pool owns one arena (hipMalloc'd once); every buffer of its contexts is a
// first-fit sub-allocation, so admitting and retiring scenes never calls hipMalloc/hipFree
// (hipFree synchronises the device) and the arena size is the admission budget.
//
// Launches.  A pool block is dft_batch steps captured once into a hipGraph.  Its kernels read a
// descriptor table (PoolDev) from device memory: the scenes, the concatenated tile list of
// their sweeps, and per kind the concatenated point groups, ADE entries and monitor points.
// The sweep kernels walk the whole tile list with a fixed, over-sized grid (the graph is
// captured once) and every kernel finds its scene by binary search over prefix offsets.  Each scene keeps its
// own device step counter: kernels skip scenes that reached their n_steps, so a scene finishing
// mid-block costs nothing further.  Changing the membership rewrites the descriptor table with
// a stream-ordered copy between blocks; the graph itself never changes.
//
// Scenes are admitted at block boundaries at step 0, so their step counters stay multiples of
// dft_batch at every boundary and the end-of-block flush replays exactly the steps since the
// previous one: a scene run in a pool accumulates the same float64 sums, in the same order, as
// the same scene run alone through oem_run.

#include "oem_internal.h"

#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#define OEM_POOL_MAX 64             // attached scenes
#define OEM_POOL_MAX_MON 256        // monitors over all attached scenes
#define OEM_POOL_SLOTS 6            // point lists in step order (points_slot)
#define OEM_POOL_STAGES 4           // pinned staging buffers for descriptor uploads
#define OEM_ARENA_ALIGN 256
// The auxiliary kernels (points, ADE, snapshots, flush) run blocks of OEM_POOL_BLK threads over
// concatenated per-scene lists whose ranges are padded to whole blocks, so every block lies in
// one scene: it finds the scene once, from a block-uniform index, and reads the scene's
// descriptor with scalar loads (a per-thread lookup would fetch it lane by lane, per entry).
#define OEM_POOL_BLK 256

static inline int pad_blk(int n) { return (n + OEM_POOL_BLK - 1) / OEM_POOL_BLK * OEM_POOL_BLK; }

struct SceneDev {
    float *E, *H;
    PsiDev ps;
    GridDev g;
    MatDev mat;
    PlaneSrcDev pw;
    DispDev disp;
    float ch, cw;
    int *stp;           // [0] steps taken, [1] steps flushed
    int n_total;
    int ny, nz;
    size_t ncell;
};

struct PtSeg {
    int scene;
    PointsDev d;
};

struct MonSeg {
    int scene;
    MonDev d;
};

struct PoolDev {
    int n_scenes, n_items, n_mon;
    int n_pt[OEM_POOL_SLOTS];
    int item_ofs[OEM_POOL_MAX + 1];
    int disp_ofs[OEM_POOL_MAX + 1];
    int pt_ofs[OEM_POOL_SLOTS][OEM_POOL_MAX + 1];
    int snap_ofs[OEM_POOL_MAX_MON + 1];
    int flush_ofs[OEM_POOL_MAX_MON + 1];
    SceneDev sc[OEM_POOL_MAX];
    PtSeg pt[OEM_POOL_SLOTS][OEM_POOL_MAX];
    MonSeg mon[OEM_POOL_MAX_MON];
};

// ------------------------------------------------------------------------------------------
// device side
// ------------------------------------------------------------------------------------------

// Largest s in [0, n) with ofs[s] <= x (ofs non-decreasing, ofs[0] = 0 <= x < ofs[n]).
__device__ __forceinline__ int find_seg(const int *ofs, int n, int x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ofs[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool scene_live(const SceneDev &S, int *n) {
    *n = S.stp[0];
    return *n < S.n_total;
}

// The grid is a multiple of the XCD count and blocks are dispatched round-robin over the XCDs,
// so block b runs on XCD b % 8 and the remap hands each XCD a contiguous run of the tile list.
__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_pool_update_h(const PoolDev *__restrict__ P)
{
    const int nit = P->n_items, ns = P->n_scenes;
    for (int w = blockIdx.x; w < nit; w += gridDim.x) {
        const int lw = xcd_remap(w, nit);
        const int s = find_seg(P->item_ofs, ns, lw);
        const SceneDev &S = P->sc[s];
        int n;
        if (!scene_live(S, &n)) continue;
        const TileCoord tc = tile_coord(S.g, lw - P->item_ofs[s]);
        if (tc.ok) update_h_tile<float>(S.H, S.E, S.ps, S.g, S.ch, S.pw, S.stp, tc);
    }
}

__global__ void __launch_bounds__(OEM_TK * OEM_TJ)
k_pool_update_e(const PoolDev *__restrict__ P)
{
    __shared__ float4 s_lut[2 * OEM_LDS_CLASSES];
    const int nit = P->n_items, ns = P->n_scenes;
    int staged = -1;
    for (int w = blockIdx.x; w < nit; w += gridDim.x) {
        const int lw = xcd_remap(w, nit);
        const int s = find_seg(P->item_ofs, ns, lw);
        const SceneDev &S = P->sc[s];
        int n;
        if (!scene_live(S, &n)) continue;             // uniform over the block
        const bool lds = lut_in_lds(S.mat);
        if (lds && s != staged) {
            __syncthreads();                          // every thread is done with the old LUT
            stage_lut(s_lut, S.mat);
            __syncthreads();
            staged = s;
        }
        const TileCoord tc = tile_coord(S.g, lw - P->item_ofs[s]);
        if (tc.ok) update_e_tile<float>(S.E, S.H, S.ps, S.g, S.mat, S.cw, S.pw, S.stp, tc, s_lut, lds);
    }
}

// Block-aligned walk over a padded list: `base` is the first entry of the block's current
// chunk (wave-uniform), the lane's entry is base + threadIdx.x.
#define OEM_POOL_CHUNKS(tot) \
    for (int base = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * OEM_POOL_BLK)); \
         base < (tot); base += gridDim.x * OEM_POOL_BLK)

__global__ void __launch_bounds__(OEM_POOL_BLK)
k_pool_points(const PoolDev *__restrict__ P, int slot)
{
    const int ns = P->n_pt[slot];
    if (ns == 0) return;
    const int *ofs = P->pt_ofs[slot];
    const int tot = ofs[ns];
    OEM_POOL_CHUNKS(tot) {
        const int g = find_seg(ofs, ns, base);
        const PtSeg &seg = P->pt[slot][g];
        const SceneDev &S = P->sc[seg.scene];
        int n;
        if (!scene_live(S, &n)) continue;
        const int t = base - ofs[g] + (int)threadIdx.x;
        if (t < seg.d.n_groups)
            point_group<float>(seg.d.e_side ? S.E : S.H, seg.d, S.mat, S.ch, S.ncell, n, t);
    }
}

// phase 0: the ADE pre of scenes at step 0; phase 1: post (last step) or fused post + pre.
__global__ void __launch_bounds__(OEM_POOL_BLK)
k_pool_disp(const PoolDev *__restrict__ P, int phase)
{
    const int ns = P->n_scenes;
    const int tot = P->disp_ofs[ns];
    OEM_POOL_CHUNKS(tot) {
        const int s = find_seg(P->disp_ofs, ns, base);
        const SceneDev &S = P->sc[s];
        int n;
        if (!scene_live(S, &n)) continue;
        int mode;
        if (phase == 0) {
            if (n != 0) continue;
            mode = 0;
        } else {
            mode = n == S.n_total - 1 ? 1 : 2;
        }
        const int e = base - P->disp_ofs[s] + (int)threadIdx.x;
        if (e < S.disp.n_entry) disp_entry<float>(S.disp, S.E, S.ncell, S.ny, S.nz, mode, e);
    }
}

__global__ void __launch_bounds__(OEM_POOL_BLK)
k_pool_snap(const PoolDev *__restrict__ P)
{
    const int nm = P->n_mon;
    if (nm == 0) return;
    const int tot = P->snap_ofs[nm];
    OEM_POOL_CHUNKS(tot) {
        const int g = find_seg(P->snap_ofs, nm, base);
        const MonSeg &M = P->mon[g];
        const SceneDev &S = P->sc[M.scene];
        int n;
        if (!scene_live(S, &n)) continue;
        const int q = base - P->snap_ofs[g] + (int)threadIdx.x;
        if (q < M.d.npts) snap_point<float>(M.d, S.E, S.H, S.ncell, S.ny, S.nz, n % M.d.T, q);
    }
}

// After the block: replay each scene's snapshots since its last flush.
__global__ void __launch_bounds__(OEM_POOL_BLK)
k_pool_flush(const PoolDev *__restrict__ P)
{
    const int nm = P->n_mon;
    if (nm == 0) return;
    const int tot = P->flush_ofs[nm];
    OEM_POOL_CHUNKS(tot) {
        const int g = find_seg(P->flush_ofs, nm, base);
        const MonSeg &M = P->mon[g];
        const SceneDev &S = P->sc[M.scene];
        const int now = S.stp[0], from = S.stp[1];
        if (now <= from) continue;
        const int r = base - P->flush_ofs[g] + (int)threadIdx.x;
        if (r >= M.d.npts * M.d.nf) continue;
        const int f = r / M.d.npts, q = r - f * M.d.npts;
        if (M.d.ncomp == 4) flush_point<float, 4>(M.d, from, now - 1, f, q);
        else flush_point<float, 6>(M.d, from, now - 1, f, q);
    }
}

__global__ void k_pool_advance(const PoolDev *__restrict__ P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P->n_scenes) return;
    int *stp = P->sc[s].stp;
    if (stp[0] < P->sc[s].n_total) stp[0] += 1;
}

__global__ void k_pool_mark_flushed(const PoolDev *__restrict__ P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P->n_scenes) return;
    int *stp = P->sc[s].stp;
    stp[1] = stp[0];
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------

struct oem_pool {
    int device = 0;
    int T = 16;
    hipStream_t stream = nullptr;
    std::mutex mu;

    // arena: free ranges by offset, live allocations by offset
    char *base = nullptr;
    size_t capacity = 0, used = 0;
    std::map<size_t, size_t> free_at;
    std::map<size_t, size_t> live;

    std::vector<oem_ctx *> members;        // attached, in attach order
    std::vector<oem_ctx *> retired;        // destroyed, memory held until retire_after is done
    int live_ctx = 0;                      // contexts of this pool not yet deleted
    bool dirty = true;

    PoolDev *d_desc = nullptr;
    PoolDev *h_stage[OEM_POOL_STAGES] = {};
    hipEvent_t stage_ev[OEM_POOL_STAGES] = {};
    bool stage_used[OEM_POOL_STAGES] = {};
    int stage_next = 0;

    hipGraphExec_t graph = nullptr;
    int grid_h = 0, grid_e = 0, grid_aux = 0, grid_pts = 0;

    int64_t launched = 0;                  // blocks enqueued
    int64_t completed = 0;                 // blocks known finished
    std::deque<std::pair<int64_t, hipEvent_t>> inflight;
};

namespace oemi {

int pool_arena_alloc(oem_pool *pool, size_t bytes, void **out) {
    const size_t need = (bytes + OEM_ARENA_ALIGN - 1) / OEM_ARENA_ALIGN * OEM_ARENA_ALIGN;
    std::lock_guard<std::mutex> lk(pool->mu);
    for (auto it = pool->free_at.begin(); it != pool->free_at.end(); ++it) {
        if (it->second < need) continue;
        const size_t off = it->first, len = it->second;
        pool->free_at.erase(it);
        if (len > need) pool->free_at[off + need] = len - need;
        pool->live[off] = need;
        pool->used += need;
        *out = pool->base + off;
        return OEM_OK;
    }
    return fail(OEM_ENOMEM, "pool arena: %zu bytes requested, %zu of %zu in use", need, pool->used,
                pool->capacity);
}

void pool_arena_free(oem_pool *pool, void *p) {
    std::lock_guard<std::mutex> lk(pool->mu);
    const size_t off = (size_t)(static_cast<char *>(p) - pool->base);
    auto lv = pool->live.find(off);
    if (lv == pool->live.end()) return;
    size_t len = lv->second, start = off;
    pool->live.erase(lv);
    pool->used -= len;
    auto nx = pool->free_at.lower_bound(off);
    if (nx != pool->free_at.end() && nx->first == off + len) {       // merge with the next range
        len += nx->second;
        nx = pool->free_at.erase(nx);
    }
    if (nx != pool->free_at.begin()) {                                // and with the previous one
        auto pv = std::prev(nx);
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            pool->free_at.erase(pv);
        }
    }
    pool->free_at[start] = len;
}

}  // namespace oemi

namespace {

// Advance `completed` over finished blocks and collect retired contexts whose memory no
// enqueued block can touch any more.  Caller holds the lock.
void reap_locked(oem_pool *pool, std::vector<oem_ctx *> &to_delete) {
    while (!pool->inflight.empty()) {
        const hipError_t q = hipEventQuery(pool->inflight.front().second);
        if (q == hipErrorNotReady) break;
        pool->completed = pool->inflight.front().first;
        (void)hipEventDestroy(pool->inflight.front().second);
        pool->inflight.pop_front();
    }
    for (size_t i = 0; i < pool->retired.size();) {
        oem_ctx *c = pool->retired[i];
        if (c->retire_after <= pool->completed) {
            to_delete.push_back(c);
            pool->retired[i] = pool->retired.back();
            pool->retired.pop_back();
        } else {
            ++i;
        }
    }
}

void delete_ctxs(oem_pool *pool, std::vector<oem_ctx *> &v) {
    for (oem_ctx *c : v) {
        delete c;                              // frees its arena ranges (takes the lock)
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->live_ctx--;
    }
    v.clear();
}

void reap(oem_pool *pool) {
    std::vector<oem_ctx *> del;
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        reap_locked(pool, del);
    }
    delete_ctxs(pool, del);
}

// Fill a host descriptor table from the members that still have steps left.
int build_desc(oem_pool *pool, PoolDev *h) {
    std::memset(h, 0, sizeof(PoolDev));
    int ns = 0, items = 0, dsp = 0, nm = 0, snap = 0, flush = 0;
    int npt[OEM_POOL_SLOTS] = {}, pto[OEM_POOL_SLOTS] = {};
    for (oem_ctx *c : pool->members) {
        if (c->step >= c->dims.n_steps) continue;
        if (ns == OEM_POOL_MAX) return fail(OEM_ENOTSUP, "more than %d running scenes", OEM_POOL_MAX);
        SceneDev &S = h->sc[ns];
        S.E = c->E;
        S.H = c->H;
        S.ps = c->psi;
        S.g = c->g;
        S.mat = c->mat;
        S.pw = c->pw;
        if (c->has_disp) S.disp = c->disp;
        S.ch = c->dims.ch;
        S.cw = c->dims.cw;
        S.stp = c->d_step;
        S.n_total = c->dims.n_steps;
        S.ny = c->dims.ny;
        S.nz = c->dims.nz;
        S.ncell = c->ncell;
        h->item_ofs[ns] = items;
        items += c->g.nblocks;
        h->disp_ofs[ns] = dsp;
        dsp += c->has_disp ? pad_blk(c->disp.n_entry) : 0;
        for (auto &pt : c->points) {
            const int slot = points_slot(pt);
            const int q = npt[slot]++;
            h->pt[slot][q].scene = ns;
            h->pt[slot][q].d = pt.d;
            h->pt_ofs[slot][q] = pto[slot];
            pto[slot] += pad_blk(pt.d.n_groups);
        }
        for (auto &m : c->mons) {
            if (nm == OEM_POOL_MAX_MON) return fail(OEM_ENOTSUP, "more than %d monitors", OEM_POOL_MAX_MON);
            h->mon[nm].scene = ns;
            h->mon[nm].d = m.d;
            h->snap_ofs[nm] = snap;
            h->flush_ofs[nm] = flush;
            snap += pad_blk(m.d.npts);
            flush += pad_blk(m.d.npts * m.d.nf);
            ++nm;
        }
        ++ns;
    }
    h->n_scenes = ns;
    h->n_items = items;
    h->item_ofs[ns] = items;
    h->disp_ofs[ns] = dsp;
    h->n_mon = nm;
    h->snap_ofs[nm] = snap;
    h->flush_ofs[nm] = flush;
    for (int s = 0; s < OEM_POOL_SLOTS; ++s) {
        h->n_pt[s] = npt[s];
        h->pt_ofs[s][npt[s]] = pto[s];
    }
    return OEM_OK;
}

int capture_graph(oem_pool *pool) {
    std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
    hipGraph_t gr = nullptr;
    HIP_OK(hipStreamBeginCapture(pool->stream, hipStreamCaptureModeThreadLocal));
    const PoolDev *P = pool->d_desc;
    const dim3 tb(OEM_TK, OEM_TJ), ab(OEM_POOL_BLK);
    for (int s = 0; s < pool->T; ++s) {
        hipLaunchKernelGGL(k_pool_update_h, dim3(pool->grid_h), tb, 0, pool->stream, P);
        for (int slot = 0; slot < 3; ++slot)
            hipLaunchKernelGGL(k_pool_points, dim3(pool->grid_pts), ab, 0, pool->stream, P, slot);
        hipLaunchKernelGGL(k_pool_disp, dim3(pool->grid_aux), ab, 0, pool->stream, P, 0);
        hipLaunchKernelGGL(k_pool_update_e, dim3(pool->grid_e), tb, 0, pool->stream, P);
        for (int slot = 3; slot < 6; ++slot)
            hipLaunchKernelGGL(k_pool_points, dim3(pool->grid_pts), ab, 0, pool->stream, P, slot);
        hipLaunchKernelGGL(k_pool_disp, dim3(pool->grid_aux), ab, 0, pool->stream, P, 1);
        hipLaunchKernelGGL(k_pool_snap, dim3(pool->grid_aux), ab, 0, pool->stream, P);
        hipLaunchKernelGGL(k_pool_advance, dim3(1), dim3(OEM_POOL_MAX), 0, pool->stream, P);
    }
    hipLaunchKernelGGL(k_pool_flush, dim3(pool->grid_aux), ab, 0, pool->stream, P);
    hipLaunchKernelGGL(k_pool_mark_flushed, dim3(1), dim3(OEM_POOL_MAX), 0, pool->stream, P);
    const hipError_t el = hipGetLastError();
    const hipError_t ec = hipStreamEndCapture(pool->stream, &gr);
    if (el != hipSuccess || ec != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return fail(OEM_EHIP, "pool graph capture: %s", hipGetErrorString(el != hipSuccess ? el : ec));
    }
    const hipError_t ei = hipGraphInstantiate(&pool->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) {
        pool->graph = nullptr;
        return fail(OEM_EHIP, "pool graph instantiate: %s", hipGetErrorString(ei));
    }
    return OEM_OK;
}

void destroy_pool(oem_pool *p) {
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (auto &f : p->inflight) (void)hipEventDestroy(f.second);
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
    for (int i = 0; i < OEM_POOL_STAGES; ++i) {
        if (p->h_stage[i]) (void)hipHostFree(p->h_stage[i]);
        if (p->stage_ev[i]) (void)hipEventDestroy(p->stage_ev[i]);
    }
    if (p->d_desc) (void)hipFree(p->d_desc);
    if (p->base) (void)hipFree(p->base);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

}  // namespace

extern "C" int oem_pool_create(int32_t device, uint64_t arena_bytes, int32_t dft_batch,
                               oem_pool **out) {
    if (!out) return fail(OEM_EINVAL, "null output");
    *out = nullptr;
    if (dft_batch < 1 || dft_batch > 64) return fail(OEM_EINVAL, "dft_batch %d outside 1..64", dft_batch);
    if (arena_bytes < (1u << 20)) return fail(OEM_EINVAL, "arena of %llu bytes", (unsigned long long)arena_bytes);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(OEM_EHIP, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(OEM_EINVAL, "device %d of %d", device, ndev);
    HIP_OK(hipSetDevice(device));
    oem_pool *p = new oem_pool();
    p->device = device;
    p->T = dft_batch;
    const size_t cap = (size_t)arena_bytes / OEM_ARENA_ALIGN * OEM_ARENA_ALIGN;
    int rc = OEM_OK;
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && !rc)
            rc = fail(e == hipErrorOutOfMemory ? OEM_ENOMEM : OEM_EHIP, "%s: %s", what, hipGetErrorString(e));
    };
    {
        std::lock_guard<std::mutex> hold(oemi::device_sync_mutex());
        hip(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (!rc) hip(hipMalloc((void **)&p->base, cap), "arena hipMalloc");
        if (!rc) hip(hipMalloc((void **)&p->d_desc, sizeof(PoolDev)), "descriptor hipMalloc");
        for (int i = 0; i < OEM_POOL_STAGES && !rc; ++i) {
            hip(hipHostMalloc((void **)&p->h_stage[i], sizeof(PoolDev), hipHostMallocDefault), "hipHostMalloc");
            if (!rc) hip(hipEventCreateWithFlags(&p->stage_ev[i], hipEventDisableTiming), "hipEventCreate");
        }
    }
    if (!rc) {
        p->capacity = cap;
        p->free_at[0] = cap;
        // an empty table, so a graph replayed before the first upload does nothing
        std::memset(p->h_stage[0], 0, sizeof(PoolDev));
        hip(hipMemcpyAsync(p->d_desc, p->h_stage[0], sizeof(PoolDev), hipMemcpyHostToDevice,
                           p->stream), "hipMemcpyAsync");
        if (!rc) hip(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    }
    if (!rc) {
        // Over-sized grids (blocks past the end of a list exit at once): with few long items
        // per block, each block's chain of dependent loads (scene lookup, step counter,
        // descriptor) runs once or twice instead of walking many items in series; 16384
        // measured 7% above a resident grid on 12 x 128^3 scenes.  A multiple of the XCD count,
        // so block b runs on XCD b % 8 and xcd_remap hands each XCD a contiguous item range.
        p->grid_h = p->grid_e = 16384;
        p->grid_aux = 16384;
        p->grid_pts = 1024;
        const char *gm = std::getenv("OEM_POOL_GRID");
        if (gm && std::atoi(gm) > 0) p->grid_h = p->grid_e = std::atoi(gm) / OEM_NXCD * OEM_NXCD;
        const char *ga = std::getenv("OEM_POOL_AUX_GRID");
        if (ga && std::atoi(ga) > 0) p->grid_aux = std::atoi(ga) / OEM_NXCD * OEM_NXCD;
        rc = capture_graph(p);
    }
    if (rc) {
        destroy_pool(p);
        return rc;
    }
    *out = p;
    return OEM_OK;
}

extern "C" int oem_pool_destroy(oem_pool *pool) {
    if (!pool) return OEM_OK;
    HIP_OK(hipSetDevice(pool->device));
    HIP_OK(hipStreamSynchronize(pool->stream));
    reap(pool);
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        if (pool->live_ctx > 0)
            return fail(OEM_ESTATE, "%d contexts of this pool are still alive", pool->live_ctx);
    }
    destroy_pool(pool);
    return OEM_OK;
}

extern "C" int oem_pool_ctx_create(oem_pool *pool, const oem_dims *d, oem_ctx **out) {
    if (!pool || !d || !out) return fail(OEM_EINVAL, "null argument");
    *out = nullptr;
    if (d->complex_fields) return fail(OEM_ENOTSUP, "the pool steps real fields only");
    if (d->dft_batch != pool->T)
        return fail(OEM_EINVAL, "dft_batch %d differs from the pool's %d", d->dft_batch, pool->T);
    reap(pool);                                   // reclaim what finished blocks released
    oem_ctx *c = new oem_ctx();
    c->pool = pool;
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->live_ctx++;
    }
    const int rc = ctx_init(c, d, pool->device);
    if (rc) {
        const std::string msg = oemi::g_err;
        std::vector<oem_ctx *> v{c};
        delete_ctxs(pool, v);
        oemi::g_err = msg;
        return rc;
    }
    *out = c;
    return OEM_OK;
}

extern "C" int oem_pool_attach(oem_pool *pool, oem_ctx *ctx) {
    if (!pool || !ctx) return fail(OEM_EINVAL, "null argument");
    if (ctx->pool != pool) return fail(OEM_EINVAL, "context belongs to another pool (or none)");
    if (!ctx->prepared) return fail(OEM_ESTATE, "oem_prepare not called");
    if (ctx->attached) return fail(OEM_ESTATE, "already attached");
    if (ctx->step != 0) return fail(OEM_ESTATE, "context already stepped %lld times", (long long)ctx->step);
    int per[OEM_POOL_SLOTS] = {};
    for (auto &pt : ctx->points)
        if (++per[points_slot(pt)] > 1) return fail(OEM_ENOTSUP, "more than one point list of one kind");
    if (!ctx->tmons.empty() || !ctx->ftmons.empty())
        return fail(OEM_ENOTSUP, "time-domain monitors run standalone only");
    HIP_OK(hipStreamSynchronize(ctx->stream));    // its uploads have landed
    std::lock_guard<std::mutex> lk(pool->mu);
    int running = 0, mons = 0;
    for (oem_ctx *c : pool->members)
        if (c->step < c->dims.n_steps) {
            ++running;
            mons += (int)c->mons.size();
        }
    if (running >= OEM_POOL_MAX) return fail(OEM_ENOTSUP, "%d scenes already running", running);
    if (mons + (int)ctx->mons.size() > OEM_POOL_MAX_MON)
        return fail(OEM_ENOTSUP, "more than %d monitors in the batch", OEM_POOL_MAX_MON);
    ctx->attached = true;
    pool->members.push_back(ctx);
    pool->dirty = true;
    return OEM_OK;
}

extern "C" int oem_pool_launch(oem_pool *pool, int32_t n_blocks) {
    if (!pool) return fail(OEM_EINVAL, "null pool");
    if (n_blocks < 0) return fail(OEM_EINVAL, "n_blocks %d", n_blocks);
    HIP_OK(hipSetDevice(pool->device));
    std::vector<oem_ctx *> del;
    int rc = OEM_OK;
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        reap_locked(pool, del);
        for (int b = 0; b < n_blocks && !rc; ++b) {
            bool any = false;
            for (oem_ctx *c : pool->members) any |= c->step < c->dims.n_steps;
            if (!any) break;
            if (pool->dirty) {
                const int si = pool->stage_next;
                pool->stage_next = (si + 1) % OEM_POOL_STAGES;
                if (pool->stage_used[si] && hipEventSynchronize(pool->stage_ev[si]) != hipSuccess) {
                    rc = fail(OEM_EHIP, "staging event");
                    break;
                }
                rc = build_desc(pool, pool->h_stage[si]);
                if (rc) break;
                if (hipMemcpyAsync(pool->d_desc, pool->h_stage[si], sizeof(PoolDev),
                                   hipMemcpyHostToDevice, pool->stream) != hipSuccess ||
                    hipEventRecord(pool->stage_ev[si], pool->stream) != hipSuccess) {
                    rc = fail(OEM_EHIP, "descriptor upload");
                    break;
                }
                pool->stage_used[si] = true;
                pool->dirty = false;
            }
            if (hipGraphLaunch(pool->graph, pool->stream) != hipSuccess) {
                rc = fail(OEM_EHIP, "pool graph launch");
                break;
            }
            hipEvent_t ev;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(ev, pool->stream) != hipSuccess) {
                rc = fail(OEM_EHIP, "block event");
                break;
            }
            const int64_t id = ++pool->launched;
            pool->inflight.emplace_back(id, ev);
            for (oem_ctx *c : pool->members) {
                if (c->step >= c->dims.n_steps) continue;
                c->step = std::min<int64_t>(c->step + pool->T, c->dims.n_steps);
                if (c->step >= c->dims.n_steps) {
                    c->done_block = id;
                    pool->dirty = true;           // drop it from the next table
                }
            }
        }
    }
    delete_ctxs(pool, del);
    return rc;
}

namespace oemi {

// Block until the pool block `id` has completed (polling: the events belong to the pool).
static int wait_block(oem_pool *pool, int64_t id) {
    for (;;) {
        std::vector<oem_ctx *> del;
        bool ok;
        {
            std::lock_guard<std::mutex> lk(pool->mu);
            reap_locked(pool, del);
            ok = pool->completed >= id;
        }
        delete_ctxs(pool, del);
        if (ok) return OEM_OK;
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

int pool_wait_done(oem_ctx *ctx) {
    if (!ctx->attached) return OEM_OK;            // never attached: only its own stream wrote
    int64_t id;
    {
        std::lock_guard<std::mutex> lk(ctx->pool->mu);
        id = ctx->done_block;
    }
    if (id < 0) return fail(OEM_ESTATE, "scene has %lld of %d steps enqueued", (long long)ctx->step,
                            ctx->dims.n_steps);
    return wait_block(ctx->pool, id);
}

int pool_retire(oem_ctx *ctx) {
    oem_pool *pool = ctx->pool;
    (void)hipStreamSynchronize(ctx->stream);
    if (!ctx->attached) {
        std::vector<oem_ctx *> v{ctx};
        delete_ctxs(pool, v);
        return OEM_OK;
    }
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        auto it = std::find(pool->members.begin(), pool->members.end(), ctx);
        if (it != pool->members.end()) pool->members.erase(it);
        // blocks enqueued so far may still read its descriptor; the next table drops it
        ctx->retire_after = pool->launched;
        ctx->step = ctx->dims.n_steps;
        pool->retired.push_back(ctx);
        pool->dirty = true;
    }
    reap(pool);
    return OEM_OK;
}

}  // namespace oemi

extern "C" int oem_pool_ctx_done(oem_pool *pool, oem_ctx *ctx, int32_t wait, int32_t *done) {
    if (!pool || !ctx || !done) return fail(OEM_EINVAL, "null argument");
    if (ctx->pool != pool || !ctx->attached) return fail(OEM_EINVAL, "context not attached to this pool");
    HIP_OK(hipSetDevice(pool->device));
    int64_t id;
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        id = ctx->done_block;
    }
    if (id < 0) {
        if (wait) return fail(OEM_ESTATE, "scene needs more blocks than were launched");
        *done = 0;
        return OEM_OK;
    }
    if (wait) TRY(oemi::wait_block(pool, id));
    reap(pool);
    std::lock_guard<std::mutex> lk(pool->mu);
    *done = pool->completed >= id ? 1 : 0;
    return OEM_OK;
}

extern "C" int oem_pool_wait(oem_pool *pool, int32_t max_pending) {
    if (!pool) return fail(OEM_EINVAL, "null pool");
    if (max_pending < 0) return fail(OEM_EINVAL, "max_pending %d", max_pending);
    HIP_OK(hipSetDevice(pool->device));
    int64_t target;
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        target = pool->launched - max_pending;
    }
    if (target > 0) TRY(oemi::wait_block(pool, target));
    return OEM_OK;
}

extern "C" int oem_pool_stats(oem_pool *pool, uint64_t *used, uint64_t *capacity,
                              int32_t *n_running, int64_t *blocks) {
    if (!pool) return fail(OEM_EINVAL, "null pool");
    reap(pool);
    std::lock_guard<std::mutex> lk(pool->mu);
    int running = 0;
    for (oem_ctx *c : pool->members) running += c->step < c->dims.n_steps;
    if (used) *used = pool->used;
    if (capacity) *capacity = pool->capacity;
    if (n_running) *n_running = running;
    if (blocks) *blocks = pool->launched;
    return OEM_OK;
}
