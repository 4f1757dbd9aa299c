// xdem_batch.hip — host side of the fused FP32c path (include/xdem_batch.h).
//
// Layout.  All scenes share one set of SoA arrays (two buffers, ping-pong).
// Scene s owns the cell keys [cell_base, cell_base + ncell) of one segmented
// grid; its particles occupy the sorted positions [p0, p0 + n), with scenes in
// cell_base order.  Storage between steps is the sorted order of the last
// sort, so a block starts with "storage sorted by (key, id)" except for newly
// added scenes, appended at the end with a sentinel key (every one of them is
// a mover of the first sort and is merged into place).
//
// A step block of n steps is
//   k_pos<adv>  sort  [k_step<adv>  sort] x (n-1)  k_step<end>
// where the last k_step leaves the end-of-step positions (no next update), so
// the state between blocks is the reference's end-of-step state.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/xdem_batch.h"
#include "../../include/xdem_io.h"
#include "xdem_batch_kernels.hpp"
#include "xdem_common.hpp"
#include "xdem_io.hpp"

using namespace xdem;
using namespace xdem::fused;

static_assert(HSLOTS == XDEM_BATCH_HISTORY_SLOTS, "history slots");
static_assert(WALL_KEY == XDEM_BATCH_WALL_KEY, "wall key");
static_assert(DRUM_KEY == XDEM_BATCH_DRUM_KEY, "drum key");

namespace {

// Slots of the displacement table, one per block of the sorted block table:
// at most n / T + scenes blocks of T >= 64 threads (each scene's last block may
// be partial).
inline size_t wdisp_words(uint64_t n, uint64_t nscene) { return n / 64 + nscene + 4; }

// A triangle-mesh object of a scene: its triangles in the object frame, sorted
// by local cell, and the local cell table (offsets filled at layout time).
struct HostObject {
    BObject ob;
    std::vector<BTri> tris;
    std::vector<uint2> cells;
};

struct HostScene {
    uint32_t id;           // user-visible scene id
    xdem_batch_scene_desc desc;
    std::vector<xdem_device_plane_info> planes;
    uint32_t n = 0;
    double rmax = 0;       // largest radius (FP32 value)
    uint32_t cell_base = 0, ncell = 0;
    uint32_t p0 = 0;       // sorted segment
    bool pending = true;   // not yet uploaded
    std::vector<double> pos, vel, omega, acc, alpha;  // staging until upload
    std::vector<uint32_t> hcount, hkey;                 // staged history (optional)
    std::vector<double> hxi;
    std::vector<BDrum> drums;                           // after the planes
    std::vector<HostObject> objects;
    uint32_t step0 = 0;    // pose-step count at admission
    uint32_t steps_before = 0;  // object steps already taken (a scene read from a checkpoint)
};

template <class T>
int32_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    XDEM_HIP(hipMalloc((void**)p, sizeof(T) * n));
    return XDEM_OK;
}

int32_t alloc_state(BState& s, size_t n) {
    XDEM_TRY(dalloc(&s.pr, n));
    XDEM_TRY(dalloc(&s.vm, n));
    XDEM_TRY(dalloc(&s.wa, n));
    XDEM_TRY(dalloc(&s.aa, n));
    XDEM_TRY(dalloc(&s.zi, n));
    XDEM_TRY(dalloc(&s.pf, n));
    return XDEM_OK;
}

void free_state(BState& s) {
    for (void* p : {(void*)s.pr, (void*)s.vm, (void*)s.wa, (void*)s.aa, (void*)s.zi, (void*)s.pf})
        if (p) (void)hipFree(p);
    s = BState{};
}

// E* of a pair (getConstant, mphysics_cuda_impl.cuh:223) and the DHS damping
// factor 4 / (1 + (pi / ln e)^2) (:239-241).
void pair_consts(const xdem_material& mi, const xdem_material& mj, const xdem_contact_pair_desc& c,
                 float& ckn, float& cvn, float& sr, float& mu) {
    double Eeq = (mi.youngs * mj.youngs) /
                 (mi.youngs * (1 - mj.poisson * mj.poisson) + mj.youngs * (1 - mi.poisson * mi.poisson));
    double beta = M_PI / std::log(c.restitution);
    ckn = (float)((4.0 / 3.0) * Eeq);
    cvn = (float)(4.0 / (1.0 + beta * beta));
    sr = (float)c.stiffness_ratio;
    mu = (float)c.friction;
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

}  // namespace

struct xdem_batch {
    uint32_t cap_np = 0, cap_scenes = 0;
    uint64_t cap_cells = 0;
    hipStream_t st = nullptr;
    BState buf[2];
    int cur = 0;                      // buffer holding the state
    // keys: skey[skcur] = sorted keys of the last sort (storage order of the
    // state); the next keys are written into skey[1 - skcur] by k_step / k_pos
    // and sorted there in place
    uint32_t* skey[2] = {nullptr, nullptr};
    int skcur = 0;
    uint32_t* perm = nullptr;         // sorted position -> storage index
    uint32_t* invperm = nullptr;      // storage index -> sorted position
    uint2* crange = nullptr;
    // sort scratch
    uint64_t* moverbits = nullptr;    // two buffers (by sort parity) of mb_words
    uint32_t* blockcnt = nullptr;     // two buffers of bc_words
    size_t mb_words = 0, bc_words = 0;
    uint32_t* counters = nullptr;     // [0],[1] movers by key parity, [2] movers of the last
                                      // sort, [3] patched (the window patch placed them),
                                      // [4],[5] scenes rebuilding every list, [6],[7] dirty tags
    uint32_t* win = nullptr;          // windows of the last two sorts (k_patch)
    uint4* mold = nullptr;            // first FEW movers: old cell size, scene, old key, first
    uint64_t* mkey = nullptr;
    uint32_t* ms = nullptr;
    uint64_t* mkey2 = nullptr;
    uint32_t* ms2 = nullptr;
    uint32_t* mplo = nullptr;         // 2 n: insertion points, then radix scratch
    // candidate lists (plane-only scenes)
    bool lists_on = false;
    bool patch_on = true;             // the window patch for few movers (else the full merge)
    float skin_frac = 0.02f;          // skin = skin_frac x cell size
    uint32_t* vl = nullptr;           // VLK entries per particle, lane-interleaved
    uint32_t* vc = nullptr;           // entries per particle
    uint16_t* dirty = nullptr;        // per cell key: tag of the sort that last marked it
    size_t dirty_bytes = 0;
    float* disp = nullptr;            // per scene (cap_scenes each)
    uint32_t* wdisp = nullptr;        // per block of k_step: its largest displacement
    uint2* d_scene_blk = nullptr;     // per scene: first and end block of the sorted table
    uint32_t* force = nullptr;
    uint32_t* rebuild = nullptr;
    // tables
    BScene* d_scenes = nullptr;
    BPlane* d_planes = nullptr;
    BDrum* d_drums = nullptr;
    BObject* d_objs = nullptr;
    BTri* d_tris = nullptr;
    uint2* d_tcells = nullptr;
    BPose* d_poses = nullptr;
    uint32_t* d_pose_ctr = nullptr;   // pose updates so far (k_poses)
    uint32_t pose_steps = 0;          // host copy of the count
    uint32_t nobj_total = 0, ndrum_total = 0;
    bool geom_active() const { return nobj_total || ndrum_total; }
    // mesh-object reactions (evaluated at the last step of each xdem_batch_step)
    bool react_on = false;
    bool react_valid = false;         // computed since the last layout change
    uint32_t react_slots = 0;         // most objects of one scene
    double* d_rpart = nullptr;        // [block][slot] 6 doubles
    size_t rpart_cap = 0;             // doubles
    uint4* d_robj = nullptr;          // per object: first and end block of its scene, slot
    double* d_react = nullptr;        // per object: force, moment
    size_t react_cap = 0;             // doubles
    BlockEntry* d_blocks_sorted = nullptr;
    BlockEntry* d_blocks_store = nullptr;
    uint32_t nblk_sorted = 0, nblk_store = 0;
    double* d_red = nullptr;          // energy partials
    unsigned long long* d_metrics = nullptr;  // contact metrics: counts, then overlap bits
    size_t metrics_cap = 0;
    // tangential-spring history (allocated with the first scene that uses it)
    float4* hxi[2] = {nullptr, nullptr};  // HSLOTS slots per particle, parity = cur
    uint32_t* hcnt[2] = {nullptr, nullptr};
    uint32_t* hover = nullptr;            // overflow counter
    bool hist_active = false;             // some scene has history on
    // scenes
    std::vector<HostScene> scenes;    // active + pending (not removed)
    uint32_t next_id = 0;
    uint32_t n_total = 0;             // particles in the arrays (incl. appended)
    bool layout_dirty = true;
    bool any_stepped = false;
    bool storage_sorted_all_new = false;
    uint32_t block_threads = 256;
    int unroll = 2;                   // k_step variant: candidates loaded together (1, 2 or 4)
    bool pipe = false;                // k_step variant: contact records loaded one ahead
    bool flat = true;                 // k_step variant: per-lane row stream (else row by row)
    bool use_graphs = true;
    std::map<std::pair<uint32_t, int>, hipGraphExec_t> graphs;
    PhaseTimer timer;
    uint32_t last_movers = 0;
    std::vector<BScene> h_scenes;
    // cell ranges of removed scenes, cleared at the next layout change (the
    // incremental sort reads the cell table around a mover's new cell)
    std::vector<std::pair<uint32_t, uint32_t>> freed_cells;
    // parity probe
    float* dbg_f = nullptr;
    uint32_t* dbg_cnt = nullptr;
    uint32_t* dbg_part = nullptr;
    uint32_t dbg_cap = 0;
    bool prepared = false;

    ~xdem_batch() {
        if (st) (void)hipStreamSynchronize(st);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        free_state(buf[0]);
        free_state(buf[1]);
        if (d_blocks_store == d_blocks_sorted) d_blocks_store = nullptr;
        for (void* p : {(void*)invperm, (void*)win, (void*)mold, (void*)vl, (void*)vc, (void*)dirty,
                        (void*)disp, (void*)wdisp, (void*)d_scene_blk, (void*)force, (void*)rebuild})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)skey[0], (void*)skey[1], (void*)perm, (void*)crange,
                        (void*)moverbits, (void*)blockcnt, (void*)counters, (void*)mkey, (void*)ms,
                        (void*)mkey2, (void*)ms2, (void*)mplo, (void*)d_scenes, (void*)d_planes,
                        (void*)d_blocks_sorted, (void*)d_blocks_store, (void*)d_red, (void*)d_metrics, (void*)dbg_f,
                        (void*)dbg_cnt, (void*)dbg_part, (void*)hxi[0], (void*)hxi[1],
                        (void*)hcnt[0], (void*)hcnt[1], (void*)hover, (void*)d_drums,
                        (void*)d_objs, (void*)d_tris, (void*)d_tcells, (void*)d_poses,
                        (void*)d_pose_ctr, (void*)d_rpart, (void*)d_robj, (void*)d_react})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }

    uint32_t nblk_rows() const { return div_up(std::max<uint32_t>(n_total, 1), ROWS_T); }

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    uint32_t key_bits() const {
        uint64_t top = 1;
        for (auto& s : scenes) top = std::max<uint64_t>(top, (uint64_t)s.cell_base + s.ncell);
        uint32_t b = 0;
        while (b < 32 && (1ull << b) < top) ++b;
        return b;
    }
    uint32_t id_bits() const {
        uint32_t mx = 1;
        for (auto& s : scenes) mx = std::max(mx, s.n);
        uint32_t b = 0;
        while (b < 32 && (1ull << b) < mx) ++b;
        return b;
    }

    // where k_step / k_pos write the keys of the positions they compute
    uint32_t* knext() const { return skey[1 - skcur]; }
    // mover marks and per-block counts of the sort of parity p (the movers of
    // the keys in skey[p]); the other parity's are cleared during a sort
    uint64_t* mbits(int p) const { return moverbits + (size_t)p * mb_words; }
    uint32_t* bcnt(int p) const { return blockcnt + (size_t)p * bc_words; }
    // both parities cleared (layout changes)
    int32_t clear_mover_marks(hipStream_t st) {
        XDEM_HIP(hipMemsetAsync(moverbits, 0, 2 * mb_words * sizeof(uint64_t), st));
        XDEM_HIP(hipMemsetAsync(blockcnt, 0, 2 * bc_words * sizeof(uint32_t), st));
        return XDEM_OK;
    }

    // Where k_step / k_pos record the movers of the keys they compute, against
    // the keys of the last sort (on = 0 when a full sort follows instead); the
    // mover counter alternates with the key buffer.
    MoverOut mover_out(bool on, const uint2* zi, const uint32_t* zperm) const {
        MoverOut m{};
        m.zi = zi;
        m.perm = zperm;
        m.key_old = skey[skcur];
        m.crange = crange;
        m.mcount = counters + (1 - skcur);
        m.mold = mold;
        m.mkey = mkey;
        m.ms = ms;
        m.mplo = mplo;
        m.moverbits = mbits(1 - skcur);
        m.blockcnt = bcnt(1 - skcur);
        m.n = n_total;
        m.ncell_cap = (uint32_t)cap_cells;
        m.on = on ? 1u : 0u;
        return m;
    }

    SortArgs sort_args(const uint2* zi) {
        SortArgs a{};
        a.n = n_total;
        a.ncell_cap = (uint32_t)cap_cells;
        a.key_old = skey[skcur];
        a.zi = zi;
        a.patched = counters + 3;
        a.crange = crange;
        a.moverbits = mbits(1 - skcur);
        a.blockcnt = bcnt(1 - skcur);
        a.msorted = counters + 2;
        a.mkey = mkey;
        a.ms = ms;
        a.mkey2 = mkey2;
        a.ms2 = ms2;
        a.mplo = mplo;
        a.perm = perm;
        a.invperm = invperm;
        a.skey = skey[1 - skcur];
        a.nblk = nblk_rows();
        a.key_bits = key_bits();
        a.id_bits = id_bits();
        return a;
    }

    PatchArgs patch_args(const uint2* zi) {
        PatchArgs a{};
        a.n = n_total;
        a.ncell_cap = (uint32_t)cap_cells;
        a.par = 1u - (uint32_t)skcur;
        a.patch_on = patch_on ? 1u : 0u;
        a.key_old = skey[skcur];
        a.skey = skey[1 - skcur];
        a.zi = zi;
        a.crange = crange;
        a.perm = perm;
        a.invperm = invperm;
        a.counters = counters;
        a.mkey = mkey;
        a.ms = ms;
        a.mplo = mplo;
        a.mold = mold;
        a.moverbits = moverbits;
        a.blockcnt = blockcnt;
        a.win = win;
        a.lists = lists_on && vl ? 1u : 0u;
        a.dirty = dirty;
        a.scenes = d_scenes;
        a.nscenes = (uint32_t)scenes.size();
        a.disp = disp;
        a.wdisp = wdisp;
        a.scene_blk = d_scene_blk;
        a.force = force;
        a.rebuild = rebuild;
        a.moverbits = mbits(1 - skcur);
        a.blockcnt = bcnt(1 - skcur);
        a.mb_clear = mbits(skcur);
        a.bc_clear = bcnt(skcur);
        a.mb_words = div_up(std::max<uint32_t>(n_total, 1), 64);
        a.bc_words = nblk_rows() + 1;
        return a;
    }
};

namespace {

// First-fit allocation of a contiguous key range for a scene's cells.
int32_t alloc_cells(xdem_batch* b, uint32_t ncell, uint32_t* base) {
    std::vector<std::pair<uint64_t, uint64_t>> used;
    for (auto& s : b->scenes)
        if (s.ncell) used.push_back({s.cell_base, (uint64_t)s.cell_base + s.ncell});
    std::sort(used.begin(), used.end());
    uint64_t at = 0;
    for (auto& u : used) {
        if (u.first >= at + ncell) break;
        at = std::max(at, u.second);
    }
    if (at + ncell > b->cap_cells || at + ncell >= 0xFFFFFFFFull) {
        set_error("cell capacity exceeded (%llu + %u > %llu)", (unsigned long long)at, ncell,
                  (unsigned long long)b->cap_cells);
        return XDEM_ERR_CAPACITY;
    }
    *base = (uint32_t)at;
    return XDEM_OK;
}

BScene make_bscene(const HostScene& h, uint32_t plane0, float skin_frac = 0.02f) {
    const xdem_batch_scene_desc& d = h.desc;
    BScene s{};
    s.wox = d.world_origin.x;
    s.woy = d.world_origin.y;
    s.woz = d.world_origin.z;
    s.cs = d.cell_size;
    s.inv_cs = 1.0 / d.cell_size;
    s.dt = d.dt;
    s.half2dt = 0.5 * d.dt * d.dt;
    s.gx = d.grid_size.x;
    s.gy = d.grid_size.y;
    s.gz = d.grid_size.z;
    s.cell_base = h.cell_base;
    s.ncell = h.ncell;
    s.p0 = h.p0;
    s.n = h.n;
    s.plane0 = plane0;
    s.nplane = (uint32_t)h.planes.size();
    s.gvx = (float)d.gravity.x;
    s.gvy = (float)d.gravity.y;
    s.gvz = (float)d.gravity.z;
    s.dtf = (float)d.dt;
    pair_consts(d.particle, d.pp.wall, d.pp, s.ckn, s.cvn, s.sr, s.mu);
    s.mass_c = (float)(d.particle.density * (4.0 / 3.0) * M_PI);
    s.rolling = d.rolling ? 1 : 0;
    s.history = d.tangential_history ? 1 : 0;
    // particle-particle cohesion (walls carry none, as on the FP64 path)
    s.cohpi = (float)(d.pp.cohesion * M_PI);
    s.rmax = h.rmax;
    s.pf_range = 2.0 * d.cell_size * std::max(d.grid_size.x, std::max(d.grid_size.y, d.grid_size.z));
    s.pf_eps = (float)std::ldexp(s.pf_range, -20);
    s.skin = (float)(skin_frac * d.cell_size);
    return s;
}

std::vector<BlockEntry> make_blocks(const std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>>& segs,
                                    uint32_t T) {
    std::vector<BlockEntry> out;
    for (auto& sg : segs) {
        uint32_t sc = sg.first, j0 = sg.second.first, n = sg.second.second;
        for (uint32_t o = 0; o < n; o += T) out.push_back(BlockEntry{sc, j0 + o, j0 + std::min(n, o + T), 0});
    }
    return out;
}

template <class T>
int32_t upload_vec(T** d, const std::vector<T>& h, hipStream_t st) {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    XDEM_TRY(dalloc(d, std::max<size_t>(h.size(), 1)));
    if (!h.empty())
        XDEM_HIP(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
    return XDEM_OK;
}

// Copy a storage segment [src, src+n) of buffer `from` to [dst, dst+n) of `to`.
int32_t copy_segment(const BState& from, const BState& to, uint32_t src, uint32_t dst, uint32_t n,
                     hipStream_t st) {
    if (!n) return XDEM_OK;
    auto cp = [&](void* d, const void* s, size_t b) -> int32_t {
        XDEM_HIP(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, st));
        return XDEM_OK;
    };
    XDEM_TRY(cp(to.pr + dst, from.pr + src, 32ull * n));
    XDEM_TRY(cp(to.vm + dst, from.vm + src, 16ull * n));
    XDEM_TRY(cp(to.wa + dst, from.wa + src, 16ull * n));
    XDEM_TRY(cp(to.aa + dst, from.aa + src, 16ull * n));
    XDEM_TRY(cp(to.zi + dst, from.zi + src, 8ull * n));
    XDEM_TRY(cp(to.pf + dst, from.pf + src, 16ull * n));
    return XDEM_OK;
}

// History buffers, zeroed (a scene without history keeps count 0).
int32_t ensure_history(xdem_batch* b) {
    if (b->hover) return XDEM_OK;
    const size_t n = b->cap_np;
    for (int k = 0; k < 2; ++k) {
        XDEM_TRY(dalloc(&b->hxi[k], n * HSLOTS));
        XDEM_TRY(dalloc(&b->hcnt[k], n));
        XDEM_HIP(hipMemsetAsync(b->hcnt[k], 0, 4 * n, b->st));
    }
    XDEM_TRY(dalloc(&b->hover, 1));
    XDEM_HIP(hipMemsetAsync(b->hover, 0, 4, b->st));
    return XDEM_OK;
}

// History of a storage segment, parity `from` -> parity `to`.
int32_t copy_hist(xdem_batch* b, int from, int to, uint32_t src, uint32_t dst, uint32_t n) {
    if (!b->hover || !n) return XDEM_OK;
    XDEM_HIP(hipMemcpyAsync(b->hxi[to] + (size_t)dst * HSLOTS, b->hxi[from] + (size_t)src * HSLOTS,
                            16ull * HSLOTS * n, hipMemcpyDeviceToDevice, b->st));
    XDEM_HIP(hipMemcpyAsync(b->hcnt[to] + dst, b->hcnt[from] + src, 4ull * n,
                            hipMemcpyDeviceToDevice, b->st));
    return XDEM_OK;
}

// Apply pending additions/removals at a block boundary (outside any graph).
int32_t apply_layout(xdem_batch* b) {
    if (!b->layout_dirty) return XDEM_OK;
    hipStream_t st = b->st;
    b->drop_graphs();
    b->hist_active = false;
    for (auto& s : b->scenes) b->hist_active |= s.desc.tangential_history != 0;
    if (b->hist_active) XDEM_TRY(ensure_history(b));
    const uint2 empty_range = make_uint2(EMPTY, 0);
    for (auto& fc : b->freed_cells) {
        hipLaunchKernelGGL(k_fill_u2, dim3(std::min<uint32_t>(4096, div_up(fc.second, 256))), dim3(256),
                           0, st, b->crange + fc.first, empty_range, (size_t)fc.second);
        XDEM_LAUNCHED();
    }
    b->freed_cells.clear();
    // 1. compaction of the scenes already on the device: they sit in the
    //    storage in cell_base order (sorted layout of the last block).
    std::vector<HostScene*> old, fresh;
    for (auto& s : b->scenes) (s.pending ? fresh : old).push_back(&s);
    std::sort(old.begin(), old.end(), [](HostScene* a, HostScene* c) { return a->p0 < c->p0; });
    BState& A = b->buf[b->cur];
    BState& B = b->buf[1 - b->cur];
    uint32_t at = 0;
    bool moved = false;
    for (HostScene* s : old) {
        if (s->p0 != at) moved = true;
        at += s->n;
    }
    uint32_t n_old = at;
    if (moved) {
        at = 0;
        for (HostScene* s : old) {
            XDEM_TRY(copy_segment(A, B, s->p0, at, s->n, st));
            XDEM_TRY(copy_hist(b, b->cur, 1 - b->cur, s->p0, at, s->n));
            XDEM_HIP(hipMemcpyAsync(b->knext() + at, b->skey[b->skcur] + s->p0, 4ull * s->n,
                                    hipMemcpyDeviceToDevice, st));
            s->p0 = at;
            at += s->n;
        }
        XDEM_TRY(copy_segment(B, A, 0, 0, n_old, st));
        XDEM_TRY(copy_hist(b, 1 - b->cur, b->cur, 0, 0, n_old));
        XDEM_HIP(hipMemcpyAsync(b->skey[b->skcur], b->knext(), 4ull * n_old, hipMemcpyDeviceToDevice,
                                st));
        // the sorted positions of the remaining scenes moved: rebuild their cell ranges
        for (HostScene* s : old) {
            hipLaunchKernelGGL(k_fill_u2, dim3(std::min<uint32_t>(4096, div_up(s->ncell, 256))),
                               dim3(256), 0, st, b->crange + s->cell_base, empty_range,
                               (size_t)s->ncell);
            XDEM_LAUNCHED();
        }
        if (n_old) {
            hipLaunchKernelGGL(k_ranges, dim3(div_up(n_old, ROWS_T)), dim3(ROWS_T), 0, st,
                               b->skey[b->skcur], b->crange, n_old, (uint32_t)b->cap_cells);
            XDEM_LAUNCHED();
        }
    }
    // 2. append new scenes (id order) with the sentinel old key
    XDEM_HIP(hipStreamSynchronize(st));
    std::vector<double4> pr;
    std::vector<float4> vm, wa, aa, pf;
    std::vector<uint2> zi;
    for (HostScene* s : fresh) {
        uint32_t n = s->n;
        s->p0 = at;  // storage position of the appended segment
        s->step0 = b->pose_steps - s->steps_before;
        const BScene bs = make_bscene(*s, 0);
        pr.resize(n);
        pf.resize(n);
        vm.resize(n);
        wa.resize(n);
        aa.assign(n, make_float4(0, 0, 0, 0));
        zi.resize(n);
        float mc = (float)(s->desc.particle.density * (4.0 / 3.0) * M_PI);
        for (uint32_t i = 0; i < n; ++i) {
            float r = (float)s->pos[4 * i + 3];
            pr[i] = make_double4(s->pos[4 * i + 0], s->pos[4 * i + 1], s->pos[4 * i + 2], (double)r);
            pf[i] = pf_of(bs, pr[i].x, pr[i].y, pr[i].z, r);
            float m = mc * (r * r * r);
            const double* v = s->vel.empty() ? nullptr : &s->vel[3 * i];
            const double* w = s->omega.empty() ? nullptr : &s->omega[3 * i];
            const double* a = s->acc.empty() ? nullptr : &s->acc[3 * i];
            const double* al = s->alpha.empty() ? nullptr : &s->alpha[3 * i];
            float alz = al ? (float)al[2] : 0.f;
            uint32_t alz_bits;
            std::memcpy(&alz_bits, &alz, 4);
            vm[i] = make_float4(v ? (float)v[0] : 0.f, v ? (float)v[1] : 0.f, v ? (float)v[2] : 0.f, m);
            wa[i] = make_float4(w ? (float)w[0] : 0.f, w ? (float)w[1] : 0.f, w ? (float)w[2] : 0.f,
                                al ? (float)al[0] : 0.f);
            if (a) aa[i] = make_float4((float)a[0], (float)a[1], (float)a[2], al ? (float)al[1] : 0.f);
            else if (al) aa[i] = make_float4(0.f, 0.f, 0.f, (float)al[1]);
            zi[i] = make_uint2(alz_bits, i);
        }
        XDEM_HIP(hipMemcpy(A.pr + at, pr.data(), 32ull * n, hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(A.vm + at, vm.data(), 16ull * n, hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(A.wa + at, wa.data(), 16ull * n, hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(A.aa + at, aa.data(), 16ull * n, hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(A.zi + at, zi.data(), 8ull * n, hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(A.pf + at, pf.data(), 16ull * n, hipMemcpyHostToDevice));
        if (b->hover) {
            // history slots in storage order (= original order for a new scene)
            std::vector<uint32_t> hc(n, 0);
            std::vector<float4> hx((size_t)n * HSLOTS, make_float4(0, 0, 0, 0));
            if (!s->hcount.empty())
                for (uint32_t i = 0; i < n; ++i) {
                    hc[i] = std::min<uint32_t>(s->hcount[i], HSLOTS);
                    for (uint32_t k = 0; k < hc[i]; ++k) {
                        size_t q = (size_t)i * HSLOTS + k;
                        uint32_t key = s->hkey[q];
                        float kf;
                        std::memcpy(&kf, &key, 4);
                        hx[q] = make_float4((float)s->hxi[3 * q], (float)s->hxi[3 * q + 1],
                                            (float)s->hxi[3 * q + 2], kf);
                    }
                }
            XDEM_HIP(hipMemcpy(b->hcnt[b->cur] + at, hc.data(), 4ull * n, hipMemcpyHostToDevice));
            XDEM_HIP(hipMemcpy(b->hxi[b->cur] + (size_t)at * HSLOTS, hx.data(), 16ull * HSLOTS * n,
                               hipMemcpyHostToDevice));
        }
        XDEM_TRY(fill_u32(b->skey[b->skcur] + at, EMPTY, n, st));
        // the scene's cells start empty
        uint2 e = make_uint2(EMPTY, 0);
        hipLaunchKernelGGL(k_fill_u2, dim3(std::min<uint32_t>(4096, div_up(s->ncell, 256))), dim3(256), 0,
                           st, b->crange + s->cell_base, e, (size_t)s->ncell);
        XDEM_LAUNCHED();
        s->pending = false;
        std::vector<double>().swap(s->pos);
        std::vector<double>().swap(s->vel);
        std::vector<double>().swap(s->omega);
        std::vector<double>().swap(s->acc);
        std::vector<double>().swap(s->alpha);
        std::vector<uint32_t>().swap(s->hcount);
        std::vector<uint32_t>().swap(s->hkey);
        std::vector<double>().swap(s->hxi);
        at += n;
    }
    b->storage_sorted_all_new = (n_old == 0);
    b->n_total = at;
    // 3. tables: the storage layout (old scenes, then the appended ones) for the
    //    first k_pos, and the sorted layout (scenes by cell_base) for everything else.
    std::vector<HostScene*> all;
    for (auto& s : b->scenes) all.push_back(&s);
    std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> store_segs, sorted_segs;
    std::vector<uint32_t> order(all.size());
    for (uint32_t i = 0; i < all.size(); ++i) order[i] = i;
    for (uint32_t i = 0; i < all.size(); ++i) store_segs.push_back({i, {all[i]->p0, all[i]->n}});
    std::sort(order.begin(), order.end(),
              [&](uint32_t x, uint32_t y) { return all[x]->cell_base < all[y]->cell_base; });
    uint32_t p = 0;
    for (uint32_t k : order) {
        all[k]->p0 = p;
        sorted_segs.push_back({k, {p, all[k]->n}});
        p += all[k]->n;
    }
    std::vector<BScene> hs;
    std::vector<BPlane> hp;
    std::vector<BDrum> hd;
    std::vector<BObject> hob;
    std::vector<BTri> htr;
    std::vector<uint2> hcl;
    std::vector<BPose> hpose;
    for (HostScene* s : all) {
        hs.push_back(make_bscene(*s, (uint32_t)hp.size(), b->skin_frac));
        float ckn, cvn, sr, mu;
        pair_consts(s->desc.particle, s->desc.wall.wall, s->desc.wall, ckn, cvn, sr, mu);
        for (auto& pl : s->planes) hp.push_back(BPlane{pl, ckn, cvn, sr, mu});
        BScene& bs = hs.back();
        bs.drum0 = (uint32_t)hd.size();
        bs.ndrum = (uint32_t)s->drums.size();
        hd.insert(hd.end(), s->drums.begin(), s->drums.end());
        bs.obj0 = (uint32_t)hob.size();
        bs.nobj = (uint32_t)s->objects.size();
        for (HostObject& ho : s->objects) {
            BObject ob = ho.ob;
            ob.tri0 = (uint32_t)htr.size();
            ob.cell0 = (uint32_t)hcl.size();
            ob.step0 = s->step0;
            htr.insert(htr.end(), ho.tris.begin(), ho.tris.end());
            hcl.insert(hcl.end(), ho.cells.begin(), ho.cells.end());
            hob.push_back(ob);
            // the current pose, for a probe before the next step (every step
            // recomputes the poses first)
            hpose.push_back(object_pose(ob, b->pose_steps - ob.step0));
        }
    }
    b->h_scenes = hs;
    b->nobj_total = (uint32_t)hob.size();
    b->ndrum_total = (uint32_t)hd.size();
    XDEM_TRY(upload_vec(&b->d_scenes, hs, st));
    XDEM_TRY(upload_vec(&b->d_planes, hp, st));
    XDEM_TRY(upload_vec(&b->d_drums, hd, st));
    XDEM_TRY(upload_vec(&b->d_objs, hob, st));
    XDEM_TRY(upload_vec(&b->d_tris, htr, st));
    XDEM_TRY(upload_vec(&b->d_tcells, hcl, st));
    XDEM_TRY(upload_vec(&b->d_poses, hpose, st));
    if (!b->d_pose_ctr) {
        XDEM_TRY(dalloc(&b->d_pose_ctr, 1));
        XDEM_HIP(hipMemsetAsync(b->d_pose_ctr, 0, 4, st));
    }
    XDEM_HIP(hipMemcpyAsync(b->d_pose_ctr, &b->pose_steps, 4, hipMemcpyHostToDevice, st));
    auto bs = make_blocks(sorted_segs, b->block_threads);
    auto bt = make_blocks(store_segs, b->block_threads);
    if (b->d_blocks_store == b->d_blocks_sorted) b->d_blocks_store = nullptr;
    XDEM_TRY(upload_vec(&b->d_blocks_sorted, bs, st));
    XDEM_TRY(upload_vec(&b->d_blocks_store, bt, st));
    b->nblk_sorted = (uint32_t)bs.size();
    b->nblk_store = (uint32_t)bt.size();
    // reaction tables: per body (every mesh object, then every drum) the block
    // range of its scene (blocks are contiguous per scene in the sorted layout)
    // and its slot in the scene (objects 0.., drums after them)
    std::vector<uint2> range(hs.size(), make_uint2(0, 0));
    std::vector<bool> seen(hs.size(), false);
    for (uint32_t L = 0; L < bs.size(); ++L) {
        const uint32_t k = bs[L].scene;
        if (!seen[k]) range[k].x = L;
        seen[k] = true;
        range[k].y = L + 1;
    }
    XDEM_REQUIRE((uint64_t)bs.size() <= wdisp_words(b->cap_np, b->cap_scenes),
                 "internal: %zu blocks exceed the wave table", bs.size());
    XDEM_TRY(upload_vec(&b->d_scene_blk, range, st));
    b->react_valid = false;
    if (b->nobj_total || b->ndrum_total) {
        std::vector<uint4> robj;
        uint32_t slots = 0;
        for (uint32_t k = 0; k < hs.size(); ++k) {
            slots = std::max(slots, hs[k].nobj + hs[k].ndrum);
            for (uint32_t q = 0; q < hs[k].nobj; ++q) robj.push_back(make_uint4(range[k].x, range[k].y, q, 0));
        }
        for (uint32_t k = 0; k < hs.size(); ++k)
            for (uint32_t d = 0; d < hs[k].ndrum; ++d)
                robj.push_back(make_uint4(range[k].x, range[k].y, hs[k].nobj + d, 0));
        b->react_slots = slots;
        XDEM_TRY(upload_vec(&b->d_robj, robj, st));
        const size_t need = (size_t)b->nblk_sorted * slots * 6;
        if (need > b->rpart_cap) {
            if (b->d_rpart) (void)hipFree(b->d_rpart);
            b->d_rpart = nullptr;
            b->rpart_cap = 0;
            XDEM_TRY(dalloc(&b->d_rpart, need));
            b->rpart_cap = need;
        }
        const size_t nbody = (size_t)b->nobj_total + b->ndrum_total;
        if (nbody * 6 > b->react_cap) {
            if (b->d_react) (void)hipFree(b->d_react);
            b->d_react = nullptr;
            b->react_cap = 0;
            XDEM_TRY(dalloc(&b->d_react, nbody * 6));
            b->react_cap = nbody * 6;
        }
    }
    // the storage moved: the sort and list state starts over
    if (b->dirty) XDEM_HIP(hipMemsetAsync(b->dirty, 0, b->dirty_bytes, st));
    XDEM_TRY(b->clear_mover_marks(st));
    hipLaunchKernelGGL(k_sort_reset, dim3(1), dim3(256), 0, st, b->patch_args(b->buf[b->cur].zi));
    XDEM_LAUNCHED();
    XDEM_HIP(hipStreamSynchronize(st));
    b->layout_dirty = false;
    return XDEM_OK;
}

// k_pos over the storage layout: Verlet position update (adv) + keys.
int32_t launch_pos(xdem_batch* b, bool adv, const BlockEntry* blocks, uint32_t nblk,
                   bool movers) {
    if (!nblk) return XDEM_OK;
    BState& A = b->buf[b->cur];
    const MoverOut mv = b->mover_out(movers, A.zi, nullptr);
    // over the contact kernel's block table the advance records its
    // displacement for the list bound (the lists carry over); otherwise every
    // list is rebuilt at the next sort
    const bool same = adv && blocks == b->d_blocks_sorted && nblk == b->nblk_sorted;
    uint32_t* wd = same ? b->wdisp : nullptr;
    uint32_t* fc = same ? nullptr : b->force;
    if (adv)
        hipLaunchKernelGGL(k_pos<true>, dim3(nblk), dim3(b->block_threads), 0, b->st, A, b->knext(),
                           b->d_scenes, blocks, mv, fc, wd);
    else
        hipLaunchKernelGGL(k_pos<false>, dim3(nblk), dim3(b->block_threads), 0, b->st, A,
                           b->knext(), b->d_scenes, blocks, mv, fc, wd);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

// The incremental (key, id) sort of keynew (storage order) against the keys of
// the previous sort; outputs perm, the new sorted keys and the cell ranges.
// The movers were recorded by the k_step / k_pos launch that wrote keynew.
int32_t run_sort(xdem_batch* b) {
    const uint32_t n = b->n_total;
    if (!n) return XDEM_OK;
    hipStream_t st = b->st;
    PhaseTimer& T = b->timer;
    const uint2* zi = b->buf[b->cur].zi;
    SortArgs a = b->sort_args(zi);
    // at most FEW movers: k_patch re-places them and the two kernels after it
    // exit; otherwise k_patch does the bookkeeping and sorts the movers (one
    // block) for the full merge (phase 3, the separate mover sort, is unused)
    XDEM_TRY(T.begin(2, st));
    hipLaunchKernelGGL(k_patch, dim3(PATCH_GRID), dim3(PATCH_T), 0, st, b->patch_args(zi), a);
    XDEM_LAUNCHED();
    XDEM_TRY(T.end(st));
    XDEM_TRY(T.begin(4, st));
    hipLaunchKernelGGL(k_merge, dim3(SWEEP_GRID + 64), dim3(ROWS_T), 0, st, a);
    XDEM_LAUNCHED();
    XDEM_TRY(T.end(st));
    b->skcur = 1 - b->skcur;
    return XDEM_OK;
}

// Full sort when every particle is new (storage in id order per scene): a
// stable radix sort of the keys with the storage index as value.
int32_t run_full_sort(xdem_batch* b) {
    const uint32_t n = b->n_total;
    hipStream_t st = b->st;
    uint32_t* out = b->knext();  // the keys k_pos wrote, sorted in place
    hipLaunchKernelGGL(k_iota, dim3(div_up(n, 256)), dim3(256), 0, st, b->perm, n);
    XDEM_LAUNCHED();
    XDEM_HIP(hipStreamSynchronize(st));  // sort_pairs may grow the library workspace
    XDEM_TRY(sort_pairs(out, b->perm, n, b->key_bits(), st));
    hipLaunchKernelGGL(k_ranges, dim3(div_up(n, ROWS_T)), dim3(ROWS_T), 0, st, out, b->crange, n,
                       (uint32_t)b->cap_cells);
    XDEM_LAUNCHED();
    hipLaunchKernelGGL(k_invert, dim3(div_up(n, 256)), dim3(256), 0, st, b->perm, b->invperm, n);
    XDEM_LAUNCHED();
    b->skcur = 1 - b->skcur;
    return XDEM_OK;
}

StepArgs step_args(xdem_batch* b, bool dbg) {
    StepArgs a{};
    a.in = b->buf[b->cur];
    a.out = b->buf[1 - b->cur];
    a.perm = b->perm;
    a.crange = b->crange;
    a.keynew = b->knext();
    a.mv = b->mover_out(true, a.in.zi, b->perm);
    a.vl = b->vl;
    a.vc = b->vc;
    a.vl_stride = b->cap_np;
    a.invperm = b->invperm;
    a.rebuild = b->rebuild;
    a.dirty = b->dirty;
    a.dtag = b->counters + 6 + b->skcur;  // the tag of the last sort (k_patch, parity skcur)
    a.wdisp = b->wdisp;
    a.force = b->force;
    a.scenes = b->d_scenes;
    a.planes = b->d_planes;
    a.drums = b->d_drums;
    a.objs = b->d_objs;
    a.tris = b->d_tris;
    a.tcells = b->d_tcells;
    a.poses = b->d_poses;
    a.blocks = b->d_blocks_sorted;
    a.nblocks = b->nblk_sorted;
    a.n = b->n_total;
    if (b->hover) {
        a.hin = b->hxi[b->cur];
        a.hcin = b->hcnt[b->cur];
        a.hout = b->hxi[1 - b->cur];
        a.hcout = b->hcnt[1 - b->cur];
        a.hoverflow = b->hover;
    }
    if (dbg) {
        a.dbg_f = b->dbg_f;
        a.dbg_cnt = b->dbg_cnt;
        a.dbg_part = b->dbg_part;
        a.dbg_max = b->dbg_cap;
    }
    return a;
}

int32_t launch_step(xdem_batch* b, bool adv) {
    if (!b->nblk_sorted) return XDEM_OK;
    if (b->nobj_total) {
        hipLaunchKernelGGL(k_poses, dim3(1), dim3(256), 0, b->st, b->d_objs, b->d_poses,
                           b->nobj_total, b->d_pose_ctr);
        XDEM_LAUNCHED();
    }
    StepArgs a = step_args(b, false);
    if (!adv && b->react_on && b->geom_active()) {
        // reactions of this (last) step, on the state and poses k_step uses
        hipLaunchKernelGGL(k_react, dim3(b->nblk_sorted), dim3(b->block_threads), 0, b->st, a,
                           b->d_rpart, b->react_slots);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(k_react_sum, dim3(b->nobj_total + b->ndrum_total), dim3(256), 0, b->st,
                           b->d_rpart, b->react_slots, b->d_robj, b->d_react);
        XDEM_LAUNCHED();
        b->react_valid = true;
    }
    XDEM_TRY(b->timer.begin(1, b->st));
    dim3 g(b->nblk_sorted), t(b->block_threads);
#define XDEM_STEP_LAUNCH(U, P, H, G)                                                          \
    do {                                                                                      \
        if (adv && b->flat)                                                                   \
            hipLaunchKernelGGL((k_step<true, false, U, P, H, G, true>), g, t, 0, b->st, a);   \
        else if (adv)                                                                         \
            hipLaunchKernelGGL((k_step<true, false, U, P, H, G, false>), g, t, 0, b->st, a);  \
        else if (b->flat)                                                                     \
            hipLaunchKernelGGL((k_step<false, false, U, P, H, G, true>), g, t, 0, b->st, a);  \
        else                                                                                  \
            hipLaunchKernelGGL((k_step<false, false, U, P, H, G, false>), g, t, 0, b->st, a); \
    } while (0)
    if (b->geom_active() && b->hist_active) {
        // drums or meshes and some scene with history (its drum springs)
        XDEM_STEP_LAUNCH(2, false, true, true);
    } else if (b->geom_active()) {
        // drums or meshes: the GEOM variants (contact records not loaded ahead)
        if (b->unroll == 1)
            XDEM_STEP_LAUNCH(1, false, false, true);
        else if (b->unroll == 4)
            XDEM_STEP_LAUNCH(4, false, false, true);
        else
            XDEM_STEP_LAUNCH(2, false, false, true);
    } else if (b->hist_active) {
        // with history the contact records are not loaded ahead
        if (b->unroll == 1)
            XDEM_STEP_LAUNCH(1, false, true, false);
        else if (b->unroll == 4)
            XDEM_STEP_LAUNCH(4, false, true, false);
        else
            XDEM_STEP_LAUNCH(2, false, true, false);
    } else if (b->flat && b->lists_on && b->vl) {
        // candidate lists (per-lane walk where a list is rebuilt); the lists
        // are kept through a block's last step (no advance) too
#define XDEM_VLIST_LAUNCH(U, P)                                                                  \
    do {                                                                                       \
        if (adv)                                                                               \
            hipLaunchKernelGGL((k_step<true, false, U, P, false, false, true, true>), g, t, 0,  \
                               b->st, a);                                                      \
        else                                                                                   \
            hipLaunchKernelGGL((k_step<false, false, U, P, false, false, true, true>), g, t, 0, \
                               b->st, a);                                                      \
    } while (0)
        const int v = b->unroll * 2 + (b->pipe ? 1 : 0);
        switch (v) {
            case 2: XDEM_VLIST_LAUNCH(1, false); break;
            case 3: XDEM_VLIST_LAUNCH(1, true); break;
            case 5: XDEM_VLIST_LAUNCH(2, true); break;
            case 8: XDEM_VLIST_LAUNCH(4, false); break;
            case 9: XDEM_VLIST_LAUNCH(4, true); break;
            default: XDEM_VLIST_LAUNCH(2, false); break;
        }
#undef XDEM_VLIST_LAUNCH
    } else {
        const int v = b->unroll * 2 + (b->pipe ? 1 : 0);
        switch (v) {
            case 2: XDEM_STEP_LAUNCH(1, false, false, false); break;
            case 3: XDEM_STEP_LAUNCH(1, true, false, false); break;
            case 5: XDEM_STEP_LAUNCH(2, true, false, false); break;
            case 8: XDEM_STEP_LAUNCH(4, false, false, false); break;
            case 9: XDEM_STEP_LAUNCH(4, true, false, false); break;
            default: XDEM_STEP_LAUNCH(2, false, false, false); break;
        }
    }
#undef XDEM_STEP_LAUNCH
    XDEM_LAUNCHED();
    XDEM_TRY(b->timer.end(b->st));
    b->cur = 1 - b->cur;
    return XDEM_OK;
}

// One block of nsteps launches (eager or under capture).
int32_t enqueue_block(xdem_batch* b, uint32_t nsteps, bool first_full_sort) {
    XDEM_TRY(b->timer.begin(0, b->st));
    XDEM_TRY(launch_pos(b, true, b->d_blocks_store, b->nblk_store, !first_full_sort));
    XDEM_TRY(b->timer.end(b->st));
    if (first_full_sort)
        XDEM_TRY(run_full_sort(b));
    else
        XDEM_TRY(run_sort(b));
    for (uint32_t k = 1; k < nsteps; ++k) {
        XDEM_TRY(launch_step(b, true));
        XDEM_TRY(run_sort(b));
    }
    XDEM_TRY(launch_step(b, false));
    return XDEM_OK;
}

}  // namespace

namespace {
// Device bytes of a batch: the arrays xdem_batch_create allocates (keep in
// step with it), the per-scene tables and block tables at their largest
// (64-thread blocks), and the history buffers ensure_history adds.
uint64_t batch_device_bytes(uint64_t n, uint64_t nscene, uint64_t ncell, bool history,
                            bool lists) {
    const uint64_t nblk = div_up((uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFull), ROWS_T);
    uint64_t b = 0;
    b += 2 * n * (sizeof(double4) + 3 * sizeof(float4) + sizeof(uint2) + sizeof(float4));  // buf[2]
    b += 4 * n * sizeof(uint32_t);                        // skey[2], perm, invperm
    b += ncell * sizeof(uint2);                           // crange
    b += 2 * (nblk * (ROWS_T / 64) + 1) * sizeof(uint64_t);    // moverbits (two parities)
    b += 2 * (nblk + 1) * sizeof(uint32_t) + 8 * sizeof(uint32_t);  // blockcnt (two), counters
    b += n * (2 * sizeof(uint64_t) + 2 * sizeof(uint32_t) + 2 * sizeof(uint32_t)) +
         FEW * sizeof(uint4);                             // movers
    b += 2 * WIN_WORDS * sizeof(uint32_t);              // windows
    b += 5 * nscene * sizeof(uint32_t);                   // disp, force, rebuild, scene blocks
    b += wdisp_words(n, nscene) * sizeof(uint32_t);       // block displacements
    b += 1024 * sizeof(double);                           // d_red
    b += nscene * sizeof(BScene) + 2 * div_up((uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFull), 64) *
                                       sizeof(BlockEntry);
    if (history) b += 2 * n * (HSLOTS * sizeof(float4) + sizeof(uint32_t)) + sizeof(uint32_t);
    if (lists) b += n * (VLK + 1) * sizeof(uint32_t) + ncell * sizeof(uint16_t);  // lists, counts, dirty tags
    return b;
}
}  // namespace

extern "C" {

uint64_t xdem_batch_device_bytes(uint32_t max_particles, uint32_t max_scenes, uint64_t max_cells,
                                 int32_t options) {
    return batch_device_bytes(max_particles, max_scenes, max_cells,
                              (options & XDEM_BATCH_FOOTPRINT_HISTORY) != 0,
                              (options & XDEM_BATCH_FOOTPRINT_LISTS) != 0);
}

int32_t xdem_batch_set_sort_patch(xdem_batch* b, int32_t on) {
    XDEM_REQUIRE(b, "null batch");
    if ((on != 0) != b->patch_on) {
        b->patch_on = on != 0;
        b->drop_graphs();
    }
    return XDEM_OK;
}

int32_t xdem_batch_list_rebuilds(xdem_batch* b, uint32_t* scenes_rebuilding) {
    XDEM_REQUIRE(b && scenes_rebuilding, "null argument");
    XDEM_HIP(hipStreamSynchronize(b->st));
    // the count of the last k_patch (its parity is the current key buffer's)
    XDEM_HIP(hipMemcpy(scenes_rebuilding, b->counters + 4 + b->skcur, 4, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int32_t xdem_batch_set_lists(xdem_batch* b, int32_t on, float skin_fraction) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(skin_fraction >= 0.0f && skin_fraction < 1.0f,
                 "skin fraction %g outside [0, 1)", (double)skin_fraction);
    if (on && !b->vl) {
        XDEM_TRY(dalloc(&b->vl, (size_t)VLK * b->cap_np));
        XDEM_TRY(dalloc(&b->vc, (size_t)b->cap_np));
        b->dirty_bytes = (size_t)b->cap_cells * sizeof(uint16_t);
        XDEM_TRY(dalloc(&b->dirty, (size_t)b->cap_cells));
        XDEM_HIP(hipMemsetAsync(b->dirty, 0, b->dirty_bytes, b->st));
        XDEM_HIP(hipStreamSynchronize(b->st));
        b->layout_dirty = true;  // the sort state restarts with the lists
    }
    if (skin_fraction > 0.0f && skin_fraction != b->skin_frac) {
        b->skin_frac = skin_fraction;
        b->layout_dirty = true;  // scene tables carry the skin
    }
    if ((on != 0) != b->lists_on) {
        b->lists_on = on != 0;
        b->drop_graphs();
        // no marks are kept while the lists are off: start from a clean state
        if (on) b->layout_dirty = true;
    }
    return XDEM_OK;
}

int32_t xdem_batch_create(uint32_t max_particles, uint32_t max_scenes, uint64_t max_cells,
                          xdem_batch** out) {
    XDEM_REQUIRE(out && max_particles && max_scenes && max_cells, "invalid capacities");
    // the contact kernel addresses its arrays with 32-bit byte offsets
    XDEM_REQUIRE(max_particles <= (1u << 27), "max_particles must be <= 2^27");
    XDEM_REQUIRE(max_cells <= (1ull << 29), "max_cells must be <= 2^29");
    *out = nullptr;
    xdem_batch* b = new xdem_batch();
    b->cap_np = max_particles;
    b->cap_scenes = max_scenes;
    b->cap_cells = max_cells;
    int32_t rc = XDEM_OK;
    auto fail = [&](int32_t r) {
        delete b;
        return r;
    };
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(XDEM_ERR_HIP);
    }
    size_t n = max_particles;
    uint32_t nblk = div_up(max_particles, ROWS_T);
    b->mb_words = (size_t)nblk * (ROWS_T / 64) + 1;
    b->bc_words = (size_t)nblk + 1;
    if ((rc = alloc_state(b->buf[0], n)) || (rc = alloc_state(b->buf[1], n)) ||
        (rc = dalloc(&b->skey[0], n)) || (rc = dalloc(&b->skey[1], n)) ||
        (rc = dalloc(&b->perm, n)) || (rc = dalloc(&b->invperm, n)) ||
        (rc = dalloc(&b->crange, (size_t)max_cells)) ||
        (rc = dalloc(&b->moverbits, 2 * ((size_t)nblk * (ROWS_T / 64) + 1))) ||
        (rc = dalloc(&b->blockcnt, 2 * ((size_t)nblk + 1))) || (rc = dalloc(&b->counters, 8)) ||
        (rc = dalloc(&b->win, 2 * WIN_WORDS)) || (rc = dalloc(&b->mold, (size_t)FEW)) ||
        (rc = dalloc(&b->mkey, n)) || (rc = dalloc(&b->ms, n)) || (rc = dalloc(&b->mkey2, n)) ||
        (rc = dalloc(&b->ms2, n)) || (rc = dalloc(&b->mplo, 2 * n)) ||
        (rc = dalloc(&b->disp, max_scenes)) ||
        (rc = dalloc(&b->wdisp, wdisp_words(max_particles, max_scenes))) ||
        (rc = dalloc(&b->force, max_scenes)) ||
        (rc = dalloc(&b->rebuild, max_scenes)) || (rc = dalloc(&b->d_red, 8 * 1024)))
        return fail(rc);
    if (hipMemset(b->counters, 0, 32) != hipSuccess ||
        hipMemset(b->win, 0, 8 * WIN_WORDS) != hipSuccess ||
        hipMemset(b->disp, 0, 4ull * max_scenes) != hipSuccess ||
        hipMemset(b->wdisp, 0, 4 * wdisp_words(max_particles, max_scenes)) != hipSuccess ||
        hipMemset(b->force, 0, 4ull * max_scenes) != hipSuccess ||
        hipMemset(b->rebuild, 0, 4ull * max_scenes) != hipSuccess ||
        hipMemset(b->moverbits, 0, 16 * ((size_t)nblk * (ROWS_T / 64) + 1)) != hipSuccess ||
        hipMemset(b->blockcnt, 0, 8 * ((size_t)nblk + 1)) != hipSuccess)
        return fail(XDEM_ERR_HIP);
    *out = b;
    return XDEM_OK;
}

int32_t xdem_batch_destroy(xdem_batch* b) {
    delete b;
    return XDEM_OK;
}

int32_t xdem_batch_add_scene(xdem_batch* b, const xdem_batch_scene_desc* d, const double* pos4,
                             const double* vel3, const double* omega3, uint32_t* scene_id) {
    return xdem_batch_add_scene_state(b, d, pos4, vel3, omega3, nullptr, nullptr, scene_id);
}

int32_t xdem_batch_add_scene_state(xdem_batch* b, const xdem_batch_scene_desc* d,
                                   const double* pos4, const double* vel3, const double* omega3,
                                   const double* acc3, const double* alpha3, uint32_t* scene_id) {
    return xdem_batch_add_scene_history(b, d, pos4, vel3, omega3, acc3, alpha3, nullptr, nullptr,
                                        nullptr, scene_id);
}

int32_t xdem_batch_add_scene_history(xdem_batch* b, const xdem_batch_scene_desc* d,
                                     const double* pos4, const double* vel3, const double* omega3,
                                     const double* acc3, const double* alpha3,
                                     const uint32_t* hist_count, const uint32_t* hist_key,
                                     const double* hist_xi, uint32_t* scene_id) {
    XDEM_REQUIRE(b && d && (d->np == 0 || pos4), "null argument");
    XDEM_REQUIRE(d->np > 0, "a scene needs particles");
    XDEM_REQUIRE(is_pow2(d->grid_size.x) && is_pow2(d->grid_size.y) && is_pow2(d->grid_size.z),
                 "grid_size must be powers of two");
    uint64_t nc = (uint64_t)d->grid_size.x * d->grid_size.y * d->grid_size.z;
    XDEM_REQUIRE(nc <= (1ull << 24), "a scene grid is limited to 2^24 cells");
    XDEM_REQUIRE(d->cell_size > 0 && d->dt > 0, "cell_size and dt must be > 0");
    XDEM_REQUIRE(d->nplane == 0 || d->planes, "null planes");
    XDEM_REQUIRE(!hist_count || (d->tangential_history && hist_key && hist_xi),
                 "history arrays need tangential_history on and all three arrays");
    XDEM_REQUIRE(d->pp.cohesion >= 0 && std::isfinite(d->pp.cohesion),
                 "cohesion must be finite and >= 0");
    if (hist_count)
        for (uint32_t i = 0; i < d->np; ++i) {
            XDEM_REQUIRE(hist_count[i] <= XDEM_BATCH_HISTORY_SLOTS, "history count %u > %d",
                         hist_count[i], XDEM_BATCH_HISTORY_SLOTS);
            for (uint32_t k = 0; k < hist_count[i]; ++k) {
                // drum keys: the scene's drums are added after it (a key of a
                // drum never added matches no contact and is dropped)
                uint32_t key = hist_key[(size_t)i * XDEM_BATCH_HISTORY_SLOTS + k];
                const bool ok = (key & DRUM_KEY) == DRUM_KEY
                                    ? (key & 3u) <= 2u && ((key & ~DRUM_KEY) >> 2) < (1u << 20)
                                : key & WALL_KEY ? (key & ~WALL_KEY) < d->nplane
                                                 : key < d->np;
                XDEM_REQUIRE(ok, "history key %#x of particle %u out of range", key, i);
            }
        }
    uint32_t active_np = 0;
    for (auto& s : b->scenes) active_np += s.n;
    if (b->scenes.size() + 1 > b->cap_scenes || (uint64_t)active_np + d->np > b->cap_np) {
        set_error("batch capacity exceeded (scenes %zu/%u, particles %u+%u/%u)", b->scenes.size(),
                  b->cap_scenes, active_np, d->np, b->cap_np);
        return XDEM_ERR_CAPACITY;
    }
    // the CFL-like requirement of the cell list: a particle may not span more than one cell
    double rmax = 0;
    for (uint32_t i = 0; i < d->np; ++i) rmax = std::max(rmax, pos4[4 * (size_t)i + 3]);
    XDEM_REQUIRE(d->cell_size >= 2.0 * (double)(float)rmax * (1.0 - 1e-7),
                 "cell_size %.9g below 2 * max radius %.9g", d->cell_size, 2 * rmax);
    HostScene h;
    h.rmax = (double)(float)rmax;
    h.id = b->next_id++;
    h.desc = *d;
    h.planes.assign(d->planes, d->planes + d->nplane);
    h.desc.planes = nullptr;
    h.n = d->np;
    h.ncell = (uint32_t)nc;
    int32_t rc = alloc_cells(b, h.ncell, &h.cell_base);
    if (rc) return rc;
    h.pos.assign(pos4, pos4 + 4 * (size_t)d->np);
    if (vel3) h.vel.assign(vel3, vel3 + 3 * (size_t)d->np);
    if (omega3) h.omega.assign(omega3, omega3 + 3 * (size_t)d->np);
    if (acc3) h.acc.assign(acc3, acc3 + 3 * (size_t)d->np);
    if (alpha3) h.alpha.assign(alpha3, alpha3 + 3 * (size_t)d->np);
    if (hist_count) {
        const size_t ns = (size_t)d->np * XDEM_BATCH_HISTORY_SLOTS;
        h.hcount.assign(hist_count, hist_count + d->np);
        h.hkey.assign(hist_key, hist_key + ns);
        h.hxi.assign(hist_xi, hist_xi + 3 * ns);
    }
    b->scenes.push_back(std::move(h));
    b->layout_dirty = true;
    b->prepared = false;
    if (scene_id) *scene_id = b->scenes.back().id;
    return XDEM_OK;
}

int32_t xdem_batch_remove_scene(xdem_batch* b, uint32_t scene_id) {
    XDEM_REQUIRE(b, "null batch");
    for (size_t i = 0; i < b->scenes.size(); ++i)
        if (b->scenes[i].id == scene_id) {
            if (!b->scenes[i].pending)
                b->freed_cells.push_back({b->scenes[i].cell_base, b->scenes[i].ncell});
            b->scenes.erase(b->scenes.begin() + i);
            b->layout_dirty = true;
            b->prepared = false;
            return XDEM_OK;
        }
    set_error("no scene %u", scene_id);
    return XDEM_ERR_INVALID;
}

int32_t xdem_batch_set_graphs(xdem_batch* b, int32_t on) {
    XDEM_REQUIRE(b, "null batch");
    b->use_graphs = on != 0;
    return XDEM_OK;
}

int32_t xdem_batch_set_block(xdem_batch* b, uint32_t threads) {
    XDEM_REQUIRE(b, "null batch");
    if (threads == 0) threads = 256;
    XDEM_REQUIRE(threads >= 64 && threads <= 256 && threads % 64 == 0,
                 "block size must be 64..256, a multiple of 64");
    if (threads != b->block_threads) {
        b->block_threads = threads;
        b->layout_dirty = true;  // rebuild block tables
    }
    return XDEM_OK;
}

int32_t xdem_batch_set_launch(xdem_batch* b, uint32_t threads, uint32_t unroll,
                              int32_t load_ahead) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(unroll == 1 || unroll == 2 || unroll == 4, "unroll must be 1, 2 or 4");
    XDEM_TRY(xdem_batch_set_block(b, threads));
    if ((int)unroll != b->unroll || (load_ahead != 0) != b->pipe) {
        b->unroll = (int)unroll;
        b->pipe = load_ahead != 0;
        b->drop_graphs();
    }
    return XDEM_OK;
}

int32_t xdem_batch_set_walk(xdem_batch* b, int32_t per_lane) {
    XDEM_REQUIRE(b, "null batch");
    if ((per_lane != 0) != b->flat) {
        b->flat = per_lane != 0;
        b->drop_graphs();
    }
    return XDEM_OK;
}

// The graph of an nsteps block from the current buffer parities: captured
// and instantiated on first use (nothing runs), then kept.
static int32_t block_graph(xdem_batch* b, uint32_t nsteps, hipGraphExec_t* out) {
    auto key = std::make_pair(nsteps, b->cur * 2 + b->skcur);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        int cur0 = b->cur, sk0 = b->skcur;
        XDEM_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int32_t rc = enqueue_block(b, nsteps, false);
        hipGraph_t gg = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->st, &gg);
        b->cur = cur0;
        b->skcur = sk0;
        if (rc) {
            if (gg) (void)hipGraphDestroy(gg);
            return rc;
        }
        XDEM_HIP(e2);
        hipGraphExec_t ex;
        hipError_t e3 = hipGraphInstantiate(&ex, gg, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gg);
        XDEM_HIP(e3);
        it = b->graphs.emplace(key, ex).first;
    }
    *out = it->second;
    return XDEM_OK;
}

int32_t xdem_batch_capture(xdem_batch* b, uint32_t nsteps) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(b->prepared, "call xdem_batch_prepare first");
    if (nsteps == 0 || !b->n_total || !b->use_graphs || b->timer.on) return XDEM_OK;
    hipGraphExec_t ex = nullptr;
    XDEM_TRY(block_graph(b, nsteps, &ex));
    if (nsteps & 1) {  // the block after this one starts from the other parities
        b->cur = 1 - b->cur;
        b->skcur = 1 - b->skcur;
        const int32_t rc = block_graph(b, nsteps, &ex);
        b->cur = 1 - b->cur;
        b->skcur = 1 - b->skcur;
        XDEM_TRY(rc);
    }
    return XDEM_OK;
}

int32_t xdem_batch_step(xdem_batch* b, uint32_t nsteps) {
    XDEM_REQUIRE(b, "null batch");
    if (nsteps == 0) return XDEM_OK;
    XDEM_TRY(apply_layout(b));
    if (!b->n_total) return XDEM_OK;
    b->prepared = false;
    bool first = b->storage_sorted_all_new;
    // The first block after a layout change runs eagerly (it uses the storage
    // table and may take the full sort); later blocks replay a captured graph.
    if (first || !b->use_graphs || b->timer.on || b->d_blocks_store != b->d_blocks_sorted) {
        XDEM_TRY(enqueue_block(b, nsteps, first));
        b->storage_sorted_all_new = false;
        // from now on the storage layout is the sorted layout
        if (b->d_blocks_store != b->d_blocks_sorted) {
            (void)hipFree(b->d_blocks_store);
            b->d_blocks_store = b->d_blocks_sorted;
            b->nblk_store = b->nblk_sorted;
        }
        b->any_stepped = true;
        if (b->nobj_total) b->pose_steps += nsteps;
        return XDEM_OK;
    }
    hipGraphExec_t ex = nullptr;
    XDEM_TRY(block_graph(b, nsteps, &ex));
    XDEM_HIP(hipGraphLaunch(ex, b->st));
    // the captured block flips buffers nsteps times (state) and nsteps times (keys)
    if (nsteps & 1) {
        b->cur = 1 - b->cur;
        b->skcur = 1 - b->skcur;
    }
    b->any_stepped = true;
    if (b->nobj_total) b->pose_steps += nsteps;
    return XDEM_OK;
}

// Autotuner: times the contact kernel (HIP events) for each candidate launch
// shape — neighbour walk (row by row, per lane, per lane with candidate lists)
// x block size 64/128/256 x candidates loaded together 1/2/4 x contact records
// loaded one ahead or not — over `steps_per_trial` real steps (the scenes
// advance) and keeps the fastest.  Returns the choice (block size + 1000 *
// unroll + 10000 * pipelined + 100000 * per-lane + 1000000 * lists) and the
// per-candidate kernel time per launch (54 entries: walk-major, then block,
// unroll, pipelined; -1 where a shape does not apply).  The step time of the
// candidate-list shapes includes their list rebuilds (the first step after a
// shape without lists rebuilds every list).
int32_t xdem_batch_autotune(xdem_batch* b, uint32_t steps_per_trial, uint32_t* chosen,
                            double* ms_per_candidate) {
    XDEM_REQUIRE(b, "null batch");
    const uint32_t cand[3] = {64, 128, 256};
    const int unr[3] = {1, 2, 4};
    if (steps_per_trial == 0) steps_per_trial = 20;
    steps_per_trial += steps_per_trial & 1u;  // even: the timed block replays the warm one's graph
    bool was_on = b->timer.on;
    XDEM_TRY(xdem_batch_enable_timing(b, 0));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XDEM_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("hipEventCreate failed");
        return XDEM_ERR_HIP;
    }
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};
    double best = 1e300;
    uint32_t best_c = b->block_threads;
    int best_u = b->unroll;
    bool best_p = b->pipe, best_f = b->flat, best_l = b->lists_on;
    XDEM_TRY(apply_layout(b));
    const bool lists_possible = !b->hist_active && !b->geom_active();
    for (int w = 0; w < 3; ++w)
    for (int i = 0; i < 3; ++i)
        for (int u = 0; u < 3; ++u)
            for (int pp = 0; pp < 2; ++pp) {
                const int slot = 18 * w + 6 * i + 2 * u + pp;
                if ((pp && (b->hist_active || b->geom_active())) || (w == 2 && !lists_possible)) {
                    // not a launch shape here
                    if (ms_per_candidate) ms_per_candidate[slot] = -1.0;
                    continue;
                }
                b->unroll = unr[u];
                b->pipe = pp != 0;
                b->flat = w >= 1;
                XDEM_TRY(xdem_batch_set_lists(b, w == 2, 0.0f));
                b->drop_graphs();
                XDEM_TRY(xdem_batch_set_block(b, cand[i]));
                XDEM_TRY(xdem_batch_step(b, 1));  // applies the new tables
                XDEM_TRY(xdem_batch_step(b, steps_per_trial));  // warm (graph captured)
                XDEM_HIP(hipEventRecord(e0, b->st));
                XDEM_TRY(xdem_batch_step(b, steps_per_trial));
                XDEM_HIP(hipEventRecord(e1, b->st));
                XDEM_HIP(hipEventSynchronize(e1));
                float el = 0.0f;
                XDEM_HIP(hipEventElapsedTime(&el, e0, e1));
                const double per = (double)el / steps_per_trial;
                if (ms_per_candidate) ms_per_candidate[slot] = per;
                if (per < best) {
                    best = per;
                    best_c = cand[i];
                    best_u = unr[u];
                    best_p = pp != 0;
                    best_f = w >= 1;
                    best_l = w == 2;
                }
            }
    b->unroll = best_u;
    b->pipe = best_p;
    b->flat = best_f;
    XDEM_TRY(xdem_batch_set_lists(b, best_l, 0.0f));
    b->drop_graphs();
    XDEM_TRY(xdem_batch_set_block(b, best_c));
    if (was_on) XDEM_TRY(xdem_batch_enable_timing(b, 1));
    if (chosen)
        *chosen = best_c + 1000u * (uint32_t)best_u + (best_p ? 10000u : 0u) +
                  (best_f ? 100000u : 0u) + (best_l ? 1000000u : 0u);
    return XDEM_OK;
}

int32_t xdem_batch_synchronize(xdem_batch* b) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_HIP(hipStreamSynchronize(b->st));
    return XDEM_OK;
}

uint32_t xdem_batch_num_particles(const xdem_batch* b) {
    uint32_t n = 0;
    if (b)
        for (auto& s : b->scenes) n += s.n;
    return n;
}

uint32_t xdem_batch_num_scenes(const xdem_batch* b) { return b ? (uint32_t)b->scenes.size() : 0; }

uint32_t xdem_batch_scene_particles(const xdem_batch* b, uint32_t scene_id) {
    if (b)
        for (auto& s : b->scenes)
            if (s.id == scene_id) return s.n;
    return 0;
}

static HostScene* find_scene(xdem_batch* b, uint32_t id) {
    for (auto& s : b->scenes)
        if (s.id == id) return &s;
    return nullptr;
}

int32_t xdem_batch_get_state(xdem_batch* b, uint32_t scene_id, double* pos4, double* vel3,
                             double* omega3, double* acc3, double* alpha3) {
    XDEM_REQUIRE(b, "null batch");
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    uint32_t n = s->n;
    if (s->pending) {  // not on the device yet: the staged inputs
        for (uint32_t i = 0; i < n; ++i) {
            if (pos4)
                for (int k = 0; k < 4; ++k)
                    pos4[4 * i + k] = k < 3 ? s->pos[4 * i + k] : (double)(float)s->pos[4 * i + 3];
            for (int k = 0; k < 3; ++k) {
                if (vel3) vel3[3 * i + k] = s->vel.empty() ? 0.0 : (double)(float)s->vel[3 * i + k];
                if (omega3) omega3[3 * i + k] = s->omega.empty() ? 0.0 : (double)(float)s->omega[3 * i + k];
                if (acc3) acc3[3 * i + k] = s->acc.empty() ? 0.0 : (double)(float)s->acc[3 * i + k];
                if (alpha3) alpha3[3 * i + k] = s->alpha.empty() ? 0.0 : (double)(float)s->alpha[3 * i + k];
            }
        }
        return XDEM_OK;
    }
    XDEM_HIP(hipStreamSynchronize(b->st));
    const BState& A = b->buf[b->cur];
    uint32_t p0 = s->p0;
    std::vector<double4> pr(n);
    std::vector<float4> vm(n), wa(n), aa(n);
    std::vector<uint2> zi(n);
    XDEM_HIP(hipMemcpy(pr.data(), A.pr + p0, 32ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(vm.data(), A.vm + p0, 16ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(wa.data(), A.wa + p0, 16ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(aa.data(), A.aa + p0, 16ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(zi.data(), A.zi + p0, 8ull * n, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t i = zi[j].y;
        XDEM_REQUIRE(i < n, "corrupt id %u at %u", i, j);
        if (pos4) {
            pos4[4 * i] = pr[j].x;
            pos4[4 * i + 1] = pr[j].y;
            pos4[4 * i + 2] = pr[j].z;
            pos4[4 * i + 3] = pr[j].w;
        }
        if (vel3) {
            vel3[3 * i] = vm[j].x;
            vel3[3 * i + 1] = vm[j].y;
            vel3[3 * i + 2] = vm[j].z;
        }
        if (omega3) {
            omega3[3 * i] = wa[j].x;
            omega3[3 * i + 1] = wa[j].y;
            omega3[3 * i + 2] = wa[j].z;
        }
        if (acc3) {
            acc3[3 * i] = aa[j].x;
            acc3[3 * i + 1] = aa[j].y;
            acc3[3 * i + 2] = aa[j].z;
        }
        if (alpha3) {
            float az;
            std::memcpy(&az, &zi[j].x, 4);
            alpha3[3 * i] = wa[j].w;
            alpha3[3 * i + 1] = aa[j].w;
            alpha3[3 * i + 2] = az;
        }
    }
    return XDEM_OK;
}

int32_t xdem_batch_get_history(xdem_batch* b, uint32_t scene_id, uint32_t* count, uint32_t* key,
                               double* xi) {
    XDEM_REQUIRE(b && count, "null argument");
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    const uint32_t n = s->n;
    const size_t ns = (size_t)n * HSLOTS;
    if (s->pending || !b->hover) {  // the staged history (or none)
        for (uint32_t i = 0; i < n; ++i) count[i] = s->hcount.empty() ? 0 : s->hcount[i];
        if (key) for (size_t q = 0; q < ns; ++q) key[q] = s->hkey.empty() ? 0 : s->hkey[q];
        if (xi) for (size_t q = 0; q < 3 * ns; ++q) xi[q] = s->hxi.empty() ? 0.0 : (double)(float)s->hxi[q];
        return XDEM_OK;
    }
    XDEM_HIP(hipStreamSynchronize(b->st));
    std::vector<uint32_t> hc(n);
    std::vector<float4> hx(ns);
    std::vector<uint2> zi(n);
    XDEM_HIP(hipMemcpy(hc.data(), b->hcnt[b->cur] + s->p0, 4ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(hx.data(), b->hxi[b->cur] + (size_t)s->p0 * HSLOTS, 16ull * ns,
                       hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(zi.data(), b->buf[b->cur].zi + s->p0, 8ull * n, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t i = zi[j].y;
        XDEM_REQUIRE(i < n, "corrupt id %u at %u", i, j);
        count[i] = std::min<uint32_t>(hc[j], HSLOTS);
        for (uint32_t k = 0; k < HSLOTS; ++k) {
            const float4 v = hx[(size_t)j * HSLOTS + k];
            const size_t q = (size_t)i * HSLOTS + k;
            bool live = k < count[i];
            if (key) {
                uint32_t kb;
                std::memcpy(&kb, &v.w, 4);
                key[q] = live ? kb : 0u;
            }
            if (xi) {
                xi[3 * q] = live ? v.x : 0.0;
                xi[3 * q + 1] = live ? v.y : 0.0;
                xi[3 * q + 2] = live ? v.z : 0.0;
            }
        }
    }
    return XDEM_OK;
}

int32_t xdem_batch_history_overflow(xdem_batch* b, uint32_t* overflow) {
    XDEM_REQUIRE(b && overflow, "null argument");
    *overflow = 0;
    if (!b->hover) return XDEM_OK;
    XDEM_HIP(hipStreamSynchronize(b->st));
    XDEM_HIP(hipMemcpy(overflow, b->hover, 4, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int32_t xdem_batch_enable_timing(xdem_batch* b, int32_t on) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_TRY(b->timer.flush());
    b->timer.reset();
    b->timer.on = on != 0;
    return XDEM_OK;
}

int32_t xdem_batch_phase_times(xdem_batch* b, double* ms6, uint32_t* count6) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_TRY(b->timer.flush());
    for (int i = 0; i < 6; ++i) {
        if (ms6) ms6[i] = b->timer.ms[i];
        if (count6) count6[i] = b->timer.count[i];
    }
    return XDEM_OK;
}

int32_t xdem_batch_last_movers(xdem_batch* b, uint32_t* movers) {
    XDEM_REQUIRE(b && movers, "null argument");
    XDEM_HIP(hipStreamSynchronize(b->st));
    XDEM_HIP(hipMemcpy(movers, b->counters + 2, 4, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

}  // extern "C"

// ------------------------------------------------------------ energy
namespace {
// Kinetic energy and momentum of one segment, FP64 accumulation, one block,
// fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_energy(BState st, uint32_t p0, uint32_t n, double* out) {
    __shared__ double s[4][1024];
    double ke = 0, mx = 0, my = 0, mz = 0;
    for (uint32_t j = threadIdx.x; j < n; j += 1024) {
        float4 vm = st.vm[p0 + j], wa = st.wa[p0 + j];
        double m = vm.w, r = st.pr[p0 + j].w;
        double vx = vm.x, vy = vm.y, vz = vm.z, wx = wa.x, wy = wa.y, wz = wa.z;
        double I = 0.4 * m * r * r;
        ke += 0.5 * m * (vx * vx + vy * vy + vz * vz) + 0.5 * I * (wx * wx + wy * wy + wz * wz);
        mx += m * vx;
        my += m * vy;
        mz += m * vz;
    }
    s[0][threadIdx.x] = ke;
    s[1][threadIdx.x] = mx;
    s[2][threadIdx.x] = my;
    s[3][threadIdx.x] = mz;
    __syncthreads();
    for (uint32_t h = 512; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 4; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}

// Contact metrics at the current positions (after xdem_batch_prepare: storage
// in sorted order, cell ranges current): per scene the particle-particle
// contacts found from each side (the reference predicate, (ri + rj) - |pj -
// pi| > 0 in FP64, over the 27 cells of the walk) and the largest such
// overlap, as the bits of a non-negative double (an integer maximum orders
// them).  A wave sums and maximises first, then one atomic per wave.
__global__ __launch_bounds__(256) void k_metrics(BState st, const uint2* crange, const BScene* scenes,
                                                 const BlockEntry* blocks, uint32_t n,
                                                 unsigned long long* cnt, unsigned long long* maxov) {
    const BlockEntry be = blocks[blockIdx.x];
    const uint32_t j = be.j0 + threadIdx.x;
    const bool live = j < be.j1 && j < n;
    uint32_t c = 0;
    double mo = 0.0;
    if (live) {
        const BScene S = scenes[be.scene];
        const double4 p = st.pr[j];
        int cx, cy, cz;
        cell_of(S, p.x, p.y, p.z, cx, cy, cz);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint2 cr = crange[key_of(S, cx + dx, cy + dy, cz + dz)];
                    if (cr.x == EMPTY) continue;
                    const uint32_t e1 = min(cr.y, n);
                    for (uint32_t q = cr.x; q < e1; ++q) {
                        if (q == j) continue;
                        const double4 o = st.pr[q];
                        const double ex = o.x - p.x, ey = o.y - p.y, ez = o.z - p.z;
                        const double cd = (p.w + o.w) - sqrt(ex * ex + ey * ey + ez * ez);
                        if (cd > 0) {
                            ++c;
                            mo = fmax(mo, cd);
                        }
                    }
                }
    }
    unsigned long long m = (unsigned long long)__double_as_longlong(mo);
    for (int o = 32; o > 0; o >>= 1) {
        c += (uint32_t)__shfl_xor((int)c, o, 64);
        const unsigned long long u = __shfl_xor(m, o, 64);
        m = u > m ? u : m;
    }
    if ((threadIdx.x & 63) == 0 && j < n && be.j0 < be.j1) {
        if (c) atomicAdd(&cnt[be.scene], (unsigned long long)c);
        if (m) atomicMax(&maxov[be.scene], m);
    }
}

// Gather the state into sorted order (perm) — used after a parity prepare so
// that storage is sorted again.
__global__ void k_gather(BState in, BState out, const uint32_t* perm, uint32_t n) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t s = perm[j];
    out.pr[j] = in.pr[s];
    out.vm[j] = in.vm[s];
    out.wa[j] = in.wa[s];
    out.aa[j] = in.aa[s];
    out.zi[j] = in.zi[s];
    out.pf[j] = in.pf[s];
}

__global__ void k_gather_hist(const float4* hin, const uint32_t* cin, float4* hout, uint32_t* cout,
                              const uint32_t* perm, uint32_t n) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t s = perm[j];
    cout[j] = cin[s];
    for (int k = 0; k < HSLOTS; ++k) hout[(size_t)j * HSLOTS + k] = hin[(size_t)s * HSLOTS + k];
}
}  // namespace

extern "C" {

int32_t xdem_batch_enable_reactions(xdem_batch* b, int32_t on) {
    XDEM_REQUIRE(b, "null batch");
    if ((on != 0) != b->react_on) {
        b->react_on = on != 0;
        b->react_valid = false;
        b->drop_graphs();
    }
    return XDEM_OK;
}

// Reactions of one scene's bodies of one kind (drums = false: mesh objects).
static int32_t body_reactions(xdem_batch* b, uint32_t scene_id, bool drums, xdem_double3* force,
                              xdem_double3* moment, uint32_t cap, uint32_t* count) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(b->react_on, "reactions are off (xdem_batch_enable_reactions)");
    uint32_t k = 0;
    while (k < b->scenes.size() && b->scenes[k].id != scene_id) ++k;
    XDEM_REQUIRE(k < b->scenes.size(), "no scene %u", scene_id);
    XDEM_REQUIRE(!b->scenes[k].pending && !b->layout_dirty && b->react_valid &&
                     k < b->h_scenes.size(),
                 "no reactions of scene %u yet (step first)", scene_id);
    const BScene& bs = b->h_scenes[k];
    // the reaction pass evaluates contacts without springs
    XDEM_REQUIRE(!drums || !bs.history || !bs.ndrum,
                 "scene %u: drum reactions are not available with tangential history", scene_id);
    const uint32_t n = drums ? bs.ndrum : bs.nobj;
    const size_t first = drums ? (size_t)b->nobj_total + bs.drum0 : (size_t)bs.obj0;
    if (count) *count = n;
    XDEM_REQUIRE(cap >= n || (!force && !moment), "capacity %u < %u bodies", cap, n);
    if (!n || (!force && !moment)) return XDEM_OK;
    std::vector<double> h((size_t)n * 6);
    XDEM_HIP(hipMemcpyAsync(h.data(), b->d_react + first * 6, 8 * h.size(), hipMemcpyDeviceToHost,
                            b->st));
    XDEM_HIP(hipStreamSynchronize(b->st));
    for (uint32_t q = 0; q < n; ++q) {
        if (force) force[q] = xdem_double3{h[6 * q], h[6 * q + 1], h[6 * q + 2]};
        if (moment) moment[q] = xdem_double3{h[6 * q + 3], h[6 * q + 4], h[6 * q + 5]};
    }
    return XDEM_OK;
}

int32_t xdem_batch_drum_reactions(xdem_batch* b, uint32_t scene_id, xdem_double3* force,
                                  xdem_double3* moment, uint32_t cap, uint32_t* ndrum) {
    return body_reactions(b, scene_id, true, force, moment, cap, ndrum);
}

int32_t xdem_batch_object_reactions(xdem_batch* b, uint32_t scene_id, xdem_double3* force,
                                    xdem_double3* moment, uint32_t cap, uint32_t* nobj) {
    return body_reactions(b, scene_id, false, force, moment, cap, nobj);
}

int32_t xdem_batch_energy(xdem_batch* b, uint32_t scene_id, double* kinetic, double momentum[3]) {
    XDEM_REQUIRE(b, "null batch");
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    XDEM_REQUIRE(!s->pending, "scene %u not on the device yet (step first)", scene_id);
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(1024), 0, b->st, b->buf[b->cur], s->p0, s->n, b->d_red);
    XDEM_LAUNCHED();
    double h[4];
    XDEM_HIP(hipMemcpyAsync(h, b->d_red, sizeof(h), hipMemcpyDeviceToHost, b->st));
    XDEM_HIP(hipStreamSynchronize(b->st));
    if (kinetic) *kinetic = h[0];
    if (momentum) {
        momentum[0] = h[1];
        momentum[1] = h[2];
        momentum[2] = h[3];
    }
    return XDEM_OK;
}

int32_t xdem_batch_contact_metrics(xdem_batch* b, uint32_t scene_id, uint64_t* contact_pairs,
                                   double* max_overlap) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(b->prepared, "call xdem_batch_prepare first");
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    XDEM_REQUIRE(!s->pending, "scene %u not on the device yet (step first)", scene_id);
    const size_t ns = std::max<size_t>(b->scenes.size(), 1);
    if (ns > b->metrics_cap) {
        if (b->d_metrics) (void)hipFree(b->d_metrics);
        b->d_metrics = nullptr;
        b->metrics_cap = 0;
        XDEM_TRY(dalloc(&b->d_metrics, 2 * ns));
        b->metrics_cap = ns;
    }
    XDEM_HIP(hipMemsetAsync(b->d_metrics, 0, 2 * ns * sizeof(unsigned long long), b->st));
    if (b->nblk_sorted) {
        hipLaunchKernelGGL(k_metrics, dim3(b->nblk_sorted), dim3(b->block_threads), 0, b->st,
                           b->buf[b->cur], b->crange, b->d_scenes, b->d_blocks_sorted, b->n_total,
                           b->d_metrics, b->d_metrics + ns);
        XDEM_LAUNCHED();
    }
    uint32_t k = 0;  // the scene's row in the device scene table
    for (; k < b->scenes.size(); ++k)
        if (&b->scenes[k] == s) break;
    unsigned long long h[2] = {0, 0};
    if (k < b->scenes.size()) {
        XDEM_HIP(hipMemcpyAsync(&h[0], b->d_metrics + k, 8, hipMemcpyDeviceToHost, b->st));
        XDEM_HIP(hipMemcpyAsync(&h[1], b->d_metrics + ns + k, 8, hipMemcpyDeviceToHost, b->st));
    }
    XDEM_HIP(hipStreamSynchronize(b->st));
    if (contact_pairs) *contact_pairs = h[0] / 2;
    if (max_overlap) {
        double v;
        std::memcpy(&v, &h[1], 8);
        *max_overlap = v;
    }
    return XDEM_OK;
}

int32_t xdem_batch_prepare(xdem_batch* b) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_TRY(apply_layout(b));
    if (!b->n_total) return XDEM_OK;
    bool first = b->storage_sorted_all_new;
    XDEM_TRY(launch_pos(b, false, b->d_blocks_store, b->nblk_store, !first));
    if (first)
        XDEM_TRY(run_full_sort(b));
    else
        XDEM_TRY(run_sort(b));
    // storage := sorted order, so the next block starts from sorted storage
    BState& A = b->buf[b->cur];
    BState& B = b->buf[1 - b->cur];
    hipLaunchKernelGGL(k_gather, dim3(div_up(b->n_total, 256)), dim3(256), 0, b->st, A, B, b->perm,
                       b->n_total);
    XDEM_LAUNCHED();
    if (b->hover) {
        hipLaunchKernelGGL(k_gather_hist, dim3(div_up(b->n_total, 256)), dim3(256), 0, b->st,
                           b->hxi[b->cur], b->hcnt[b->cur], b->hxi[1 - b->cur],
                           b->hcnt[1 - b->cur], b->perm, b->n_total);
        XDEM_LAUNCHED();
    }
    b->cur = 1 - b->cur;
    hipLaunchKernelGGL(k_iota, dim3(div_up(b->n_total, 256)), dim3(256), 0, b->st, b->perm, b->n_total);
    XDEM_LAUNCHED();
    hipLaunchKernelGGL(k_iota, dim3(div_up(b->n_total, 256)), dim3(256), 0, b->st, b->invperm,
                       b->n_total);
    XDEM_LAUNCHED();
    // storage moved: the window sets and the lists start over
    hipLaunchKernelGGL(k_sort_reset, dim3(1), dim3(256), 0, b->st, b->patch_args(b->buf[b->cur].zi));
    XDEM_LAUNCHED();
    b->storage_sorted_all_new = false;
    if (b->d_blocks_store != b->d_blocks_sorted) {
        (void)hipFree(b->d_blocks_store);
        b->d_blocks_store = b->d_blocks_sorted;
        b->nblk_store = b->nblk_sorted;
    }
    b->drop_graphs();
    b->prepared = true;
    XDEM_HIP(hipStreamSynchronize(b->st));
    return XDEM_OK;
}

int32_t xdem_batch_probe(xdem_batch* b, uint32_t scene_id, double* force3, double* moment3,
                         uint32_t max_per, uint32_t* counts, uint32_t* partners,
                         uint32_t* sorted_keys, uint32_t* sorted_ids, uint32_t* cstart,
                         uint32_t* cend) {
    XDEM_REQUIRE(b, "null batch");
    XDEM_REQUIRE(b->prepared, "call xdem_batch_prepare first");
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    uint32_t mp = std::max<uint32_t>(max_per, 1);
    if (b->dbg_cap != mp || !b->dbg_f) {
        if (b->dbg_f) (void)hipFree(b->dbg_f);
        if (b->dbg_cnt) (void)hipFree(b->dbg_cnt);
        if (b->dbg_part) (void)hipFree(b->dbg_part);
        b->dbg_f = nullptr;
        b->dbg_cnt = nullptr;
        b->dbg_part = nullptr;
        XDEM_TRY(dalloc(&b->dbg_f, 6ull * b->cap_np));
        XDEM_TRY(dalloc(&b->dbg_cnt, (size_t)b->cap_np));
        XDEM_TRY(dalloc(&b->dbg_part, (size_t)b->cap_np * mp));
        b->dbg_cap = mp;
    }
    StepArgs a = step_args(b, true);  // no advance: keynew is not written
    if (b->geom_active())
        hipLaunchKernelGGL((k_step<false, true, 2, false, false, true>), dim3(b->nblk_sorted),
                           dim3(b->block_threads), 0, b->st, a);
    else
        hipLaunchKernelGGL((k_step<false, true, 2, false, false>), dim3(b->nblk_sorted),
                           dim3(b->block_threads), 0, b->st, a);
    XDEM_LAUNCHED();
    XDEM_HIP(hipStreamSynchronize(b->st));
    uint32_t n = s->n, p0 = s->p0;
    std::vector<float> f(6ull * n);
    std::vector<uint32_t> cnt(n), part((size_t)n * mp), sk(n);
    std::vector<uint2> zi(n);
    XDEM_HIP(hipMemcpy(f.data(), b->dbg_f + 6ull * p0, 24ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(cnt.data(), b->dbg_cnt + p0, 4ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(part.data(), b->dbg_part + (size_t)p0 * mp, 4ull * n * mp, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(zi.data(), b->buf[b->cur].zi + p0, 8ull * n, hipMemcpyDeviceToHost));
    XDEM_HIP(hipMemcpy(sk.data(), b->skey[b->skcur] + p0, 4ull * n, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t i = zi[j].y;
        XDEM_REQUIRE(i < n, "corrupt id");
        for (int k = 0; k < 3; ++k) {
            if (force3) force3[3 * i + k] = f[6 * j + k];
            if (moment3) moment3[3 * i + k] = f[6 * j + 3 + k];
        }
        if (counts) counts[i] = cnt[j];
        if (partners)
            for (uint32_t c = 0; c < mp; ++c) partners[(size_t)i * mp + c] = part[(size_t)j * mp + c];
        if (sorted_keys) sorted_keys[j] = sk[j] - s->cell_base;
        if (sorted_ids) sorted_ids[j] = i;
    }
    if (cstart || cend) {
        std::vector<uint2> cr(s->ncell);
        XDEM_HIP(hipMemcpy(cr.data(), b->crange + s->cell_base, 8ull * s->ncell, hipMemcpyDeviceToHost));
        for (uint32_t c = 0; c < s->ncell; ++c) {
            bool empty = cr[c].x == EMPTY;
            if (cstart) cstart[c] = empty ? EMPTY : cr[c].x - p0;
            if (cend) cend[c] = empty ? 0u : cr[c].y - p0;
        }
    }
    return XDEM_OK;
}

namespace {

// numeric::utility::calculate_center_of_triangle (source/numeric_utility.cpp:20-45),
// the circumcentre updatePolygonObjectData also uses.
d3 circumcenter_host(d3 P, d3 Q, d3 R) {
    d3 V = Q - P, W = R - P;
    d3 N = cross(V, W);
    N = N / std::sqrt(dot(N, N));
    d3 M1 = 0.5 * (Q + P), M2 = 0.5 * (R + P);
    d3 D1 = cross(N, V), D2 = cross(N, W);
    double t = 0.0;
    if (std::fabs(D1.x * D2.y - D1.y * D2.x) > 1E-13)
        t = (D2.x * (M1.y - M2.y)) / (D1.x * D2.y - D1.y * D2.x) -
            (D2.y * (M1.x - M2.x)) / (D1.x * D2.y - D1.y * D2.x);
    else if (std::fabs(D1.x * D2.z - D1.z * D2.x) > 1E-13)
        t = (D2.x * (M1.z - M2.z)) / (D1.x * D2.z - D1.z * D2.x) -
            (D2.z * (M1.x - M2.x)) / (D1.x * D2.z - D1.z * D2.x);
    else if (std::fabs(D1.y * D2.z - D1.z * D2.y) > 1E-13)
        t = (D2.y * (M1.z - M2.z)) / (D1.y * D2.z - D1.z * D2.y) -
            (D2.z * (M1.y - M2.y)) / (D1.y * D2.z - D1.z * D2.y);
    return M1 + t * D1;
}

// mesh: a polygon object is being added (no tangential history with meshes;
// drums carry their springs in the particles' history)
int32_t pending_scene(xdem_batch* b, uint32_t scene_id, bool mesh, HostScene** out) {
    HostScene* s = find_scene(b, scene_id);
    XDEM_REQUIRE(s, "no scene %u", scene_id);
    XDEM_REQUIRE(s->pending, "scene %u has stepped: add its walls before its first step", scene_id);
    XDEM_REQUIRE(!mesh || !s->desc.tangential_history,
                 "scene %u: tangential history is not available with polygon objects", scene_id);
    *out = s;
    return XDEM_OK;
}

}  // namespace

int32_t xdem_batch_add_drum(xdem_batch* b, uint32_t scene_id, const xdem_drum_info* drum,
                            const xdem_contact_pair_desc* wall) {
    XDEM_REQUIRE(b && drum && wall, "null argument");
    HostScene* s = nullptr;
    XDEM_TRY(pending_scene(b, scene_id, false, &s));
    XDEM_REQUIRE(s->drums.size() < (1u << 20), "scene %u: too many drums", scene_id);
    const xdem_double3& a = drum->axis;
    double an = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    XDEM_REQUIRE(std::fabs(an - 1.0) < 1e-12, "drum axis must be a unit vector (|axis| = %.17g)", an);
    XDEM_REQUIRE(drum->radius > 0 && drum->half_length > 0, "drum radius and half length must be > 0");
    BDrum bd{};
    bd.d = *drum;
    pair_consts(s->desc.particle, wall->wall, *wall, bd.ckn, bd.cvn, bd.sr, bd.mu);
    s->drums.push_back(bd);
    b->layout_dirty = true;
    return XDEM_OK;
}

int32_t xdem_batch_add_polygon_object(xdem_batch* b, uint32_t scene_id,
                                      const xdem_device_polygon_info* tris, uint32_t ntri,
                                      const xdem_double3* origin, const xdem_double4* ep,
                                      const xdem_motion* motion,
                                      const xdem_contact_pair_desc* wall) {
    XDEM_REQUIRE(b && tris && ntri && origin && ep && wall, "null argument");
    HostScene* s = nullptr;
    XDEM_TRY(pending_scene(b, scene_id, true, &s));
    HostObject ho{};
    BObject& ob = ho.ob;
    ob.origin0 = *origin;
    ob.ep0 = *ep;
    ob.vel = motion ? motion->velocity : xdem_double3{0, 0, 0};
    ob.omega = motion ? motion->omega : xdem_double3{0, 0, 0};
    ob.dt = s->desc.dt;
    // the motion's active steps: k dt >= start_time and, when end_time >
    // start_time, k dt <= end_time (k = 1, 2, ...; the same products as the
    // simulation object's step time)
    ob.k_on = 1;
    ob.k_off = 0xFFFFFFFFu;
    if (motion) {
        const double dt = ob.dt, t0 = motion->start_time, t1 = motion->end_time;
        XDEM_REQUIRE(std::isfinite(t0) && std::isfinite(t1) && t0 / dt < 4.0e9 && t1 / dt < 4.0e9,
                     "motion times must be finite and below 4e9 steps");
        uint64_t k = (uint64_t)std::max(1.0, std::floor(t0 / dt));
        while ((double)k * dt < t0) ++k;
        while (k > 1 && (double)(k - 1) * dt >= t0) --k;
        ob.k_on = (uint32_t)k;
        if (t1 > t0) {
            uint64_t e = (uint64_t)std::max(0.0, std::floor(t1 / dt));
            while (e > 0 && (double)e * dt > t1) --e;
            while ((double)(e + 1) * dt <= t1) ++e;
            ob.k_off = (uint32_t)e;  // below k_on: never active
        }
    }
    ob.cs = s->desc.cell_size;
    pair_consts(s->desc.particle, wall->wall, *wall, ob.ckn, ob.cvn, ob.sr, ob.mu);
    // local vertices A(ep0)^T (P - origin0) (toLocal, mphysics_cuda_impl.cuh:640-645)
    const xdem_double4& e = *ep;
    const d3 r0 = mk3(2.0 * (e.x * e.x + e.y * e.y - 0.5), 2.0 * (e.y * e.z - e.x * e.w),
                      2.0 * (e.y * e.w + e.x * e.z));
    const d3 r1 = mk3(2.0 * (e.y * e.z + e.x * e.w), 2.0 * (e.x * e.x + e.z * e.z - 0.5),
                      2.0 * (e.z * e.w - e.x * e.y));
    const d3 r2 = mk3(2.0 * (e.y * e.w - e.x * e.z), 2.0 * (e.z * e.w + e.x * e.y),
                      2.0 * (e.x * e.x + e.w * e.w - 0.5));
    auto to_local = [&](const xdem_double3& v) {
        d3 d = mk3(v.x - origin->x, v.y - origin->y, v.z - origin->z);
        return mk3(r0.x * d.x + r1.x * d.y + r2.x * d.z, r0.y * d.x + r1.y * d.y + r2.y * d.z,
                   r0.z * d.x + r1.z * d.y + r2.z * d.z);
    };
    std::vector<BTri> lt(ntri);
    std::vector<d3> cc(ntri);
    double rc_max = 0.0;
    d3 lo = mk3(1e300, 1e300, 1e300), hi = mk3(-1e300, -1e300, -1e300);
    for (uint32_t t = 0; t < ntri; ++t) {
        d3 P = to_local(tris[t].P), Q = to_local(tris[t].Q), R = to_local(tris[t].R);
        d3 n = -cross(Q - P, R - P);
        double nl = std::sqrt(dot(n, n));
        XDEM_REQUIRE(nl > 0, "triangle %u is degenerate", t);
        n = n / nl;
        cc[t] = circumcenter_host(P, Q, R);
        d3 dp = cc[t] - P, dq = cc[t] - Q, dr = cc[t] - R;
        const double rc = std::sqrt(std::max(dot(dp, dp), std::max(dot(dq, dq), dot(dr, dr))));
        rc_max = std::max(rc_max, std::sqrt(dot(dp, dp)));
        lo = mk3(std::min(lo.x, cc[t].x), std::min(lo.y, cc[t].y), std::min(lo.z, cc[t].z));
        hi = mk3(std::max(hi.x, cc[t].x), std::max(hi.y, cc[t].y), std::max(hi.z, cc[t].z));
        lt[t] = BTri{cc[t].x, cc[t].y, cc[t].z, rc, P.x, P.y, P.z, Q.x, Q.y, Q.z, R.x, R.y, R.z,
                     (float)n.x, (float)n.y, (float)n.z, 0.0f};
    }
    XDEM_REQUIRE(ob.cs >= s->rmax + rc_max,
                 "cell size %.9g is below the largest radius %.9g plus the largest triangle "
                 "circumradius %.9g: the triangle search would miss contacts",
                 ob.cs, s->rmax, rc_max);
    ob.lox = lo.x - ob.cs;
    ob.loy = lo.y - ob.cs;
    ob.loz = lo.z - ob.cs;
    ob.gx = (uint32_t)std::floor((hi.x - ob.lox) / ob.cs) + 2;
    ob.gy = (uint32_t)std::floor((hi.y - ob.loy) / ob.cs) + 2;
    ob.gz = (uint32_t)std::floor((hi.z - ob.loz) / ob.cs) + 2;
    uint64_t ncell = (uint64_t)ob.gx * ob.gy * ob.gz;
    XDEM_REQUIRE(ncell <= (1ull << 26), "object grid of %llu cells (extent / cell size too large)",
                 (unsigned long long)ncell);
    std::vector<std::pair<uint32_t, uint32_t>> key(ntri);
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t x = (uint32_t)std::floor((cc[t].x - ob.lox) / ob.cs);
        uint32_t y = (uint32_t)std::floor((cc[t].y - ob.loy) / ob.cs);
        uint32_t z = (uint32_t)std::floor((cc[t].z - ob.loz) / ob.cs);
        key[t] = {(z * ob.gy + y) * ob.gx + x, t};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& c) {
                         return a.first < c.first;
                     });
    ho.cells.assign(ncell, make_uint2(0, 0));
    ho.tris.resize(ntri);
    for (uint32_t k = 0; k < ntri; ++k) {
        ho.tris[k] = lt[key[k].second];
        uint2& c = ho.cells[key[k].first];
        if (k == 0 || key[k - 1].first != key[k].first) c.x = k;
        c.y = k + 1;
    }
    ob.ntri = ntri;
    s->objects.push_back(std::move(ho));
    b->layout_dirty = true;
    return XDEM_OK;
}

// Payload: scene count; per scene its description (planes pointer cleared),
// planes, the five state arrays in original particle order, and with
// tangential history the slot counts, keys and springs.
int32_t xdem_batch_save_checkpoint(xdem_batch* b, const char* path) {
    XDEM_REQUIRE(b, "null batch");
    io::Writer w;
    w.put((uint32_t)b->scenes.size());
    std::vector<uint32_t> ids;
    for (auto& s : b->scenes) ids.push_back(s.id);
    for (uint32_t id : ids) {
        HostScene* s = find_scene(b, id);
        const size_t n = s->n, ns = n * XDEM_BATCH_HISTORY_SLOTS;
        xdem_batch_scene_desc d;
        std::memset(&d, 0, sizeof(d));
        d = s->desc;
        d.planes = nullptr;
        d.nplane = (uint32_t)s->planes.size();
        w.put(d);
        w.arr(s->planes.data(), s->planes.size());
        std::vector<double> pos(4 * n), vel(3 * n), om(3 * n), acc(3 * n), al(3 * n);
        XDEM_TRY(xdem_batch_get_state(b, id, pos.data(), vel.data(), om.data(), acc.data(),
                                      al.data()));
        w.arr(pos.data(), pos.size());
        w.arr(vel.data(), vel.size());
        w.arr(om.data(), om.size());
        w.arr(acc.data(), acc.size());
        w.arr(al.data(), al.size());
        if (d.tangential_history) {
            std::vector<uint32_t> cnt(n), key(ns);
            std::vector<double> xi(3 * ns);
            XDEM_TRY(xdem_batch_get_history(b, id, cnt.data(), key.data(), xi.data()));
            w.arr(cnt.data(), cnt.size());
            w.arr(key.data(), key.size());
            w.arr(xi.data(), xi.size());
        }
        // drums, polygon objects and the object steps already taken
        w.put((uint32_t)s->drums.size());
        w.arr(s->drums.data(), s->drums.size());
        w.put((uint32_t)s->objects.size());
        for (const HostObject& ho : s->objects) {
            w.put(ho.ob);
            w.put((uint32_t)ho.tris.size());
            w.arr(ho.tris.data(), ho.tris.size());
            w.put((uint32_t)ho.cells.size());
            w.arr(ho.cells.data(), ho.cells.size());
        }
        w.put(s->pending ? s->steps_before : b->pose_steps - s->step0);
    }
    return io::write_checkpoint(path, io::CKPT_BATCH, w);
}

int32_t xdem_batch_load_checkpoint(xdem_batch* b, const char* path, uint32_t* scene_ids,
                                   uint32_t cap, uint32_t* nscenes) {
    XDEM_REQUIRE(b, "null batch");
    std::vector<uint8_t> file;
    io::Reader r;
    XDEM_TRY(io::read_checkpoint(path, io::CKPT_BATCH, file, r, nullptr));
    uint32_t ns_total;
    XDEM_REQUIRE(r.get(ns_total), "%s: short header", path);
    XDEM_REQUIRE(!scene_ids || cap >= ns_total, "%s: %u scenes, room for %u ids", path, ns_total,
                 cap);
    for (uint32_t k = 0; k < ns_total; ++k) {
        xdem_batch_scene_desc d;
        XDEM_REQUIRE(r.get(d), "%s: truncated scene %u", path, k);
        XDEM_REQUIRE(d.np > 0 && d.nplane <= (1u << 20), "%s: bad scene %u", path, k);
        const size_t n = d.np, ns = n * XDEM_BATCH_HISTORY_SLOTS;
        std::vector<xdem_device_plane_info> planes(d.nplane);
        std::vector<double> pos(4 * n), vel(3 * n), om(3 * n), acc(3 * n), al(3 * n);
        bool ok = r.arr(planes.data(), planes.size()) && r.arr(pos.data(), pos.size()) &&
                  r.arr(vel.data(), vel.size()) && r.arr(om.data(), om.size()) &&
                  r.arr(acc.data(), acc.size()) && r.arr(al.data(), al.size());
        std::vector<uint32_t> cnt, key;
        std::vector<double> xi;
        if (ok && d.tangential_history) {
            cnt.resize(n);
            key.resize(ns);
            xi.resize(3 * ns);
            ok = r.arr(cnt.data(), n) && r.arr(key.data(), ns) && r.arr(xi.data(), 3 * ns);
        }
        uint32_t nd = 0, no = 0, before = 0;
        std::vector<BDrum> drums;
        std::vector<HostObject> objs;
        ok = ok && r.get(nd) && nd <= (1u << 16);
        if (ok) {
            drums.resize(nd);
            ok = r.arr(drums.data(), nd) && r.get(no) && no <= (1u << 16);
        }
        for (uint32_t q = 0; ok && q < no; ++q) {
            HostObject ho{};
            uint32_t nt = 0, nc = 0;
            ok = r.get(ho.ob) && r.get(nt) && nt <= (1u << 28);
            if (ok) {
                ho.tris.resize(nt);
                ok = r.arr(ho.tris.data(), nt) && r.get(nc) && nc <= (1u << 26);
            }
            if (ok) {
                ho.cells.resize(nc);
                ok = r.arr(ho.cells.data(), nc) &&
                     nc == (uint64_t)ho.ob.gx * ho.ob.gy * ho.ob.gz && nt == ho.ob.ntri;
            }
            if (ok)
                for (const uint2& c : ho.cells) ok = ok && c.x <= c.y && c.y <= nt;
            objs.push_back(std::move(ho));
        }
        ok = ok && r.get(before);
        XDEM_REQUIRE(ok, "%s: truncated or inconsistent scene %u", path, k);
        d.planes = planes.empty() ? nullptr : planes.data();
        uint32_t id = 0;
        XDEM_TRY(xdem_batch_add_scene_history(
            b, &d, pos.data(), vel.data(), om.data(), acc.data(), al.data(),
            d.tangential_history ? cnt.data() : nullptr, d.tangential_history ? key.data() : nullptr,
            d.tangential_history ? xi.data() : nullptr, &id));
        HostScene* hs = find_scene(b, id);
        hs->drums = std::move(drums);
        hs->objects = std::move(objs);
        hs->steps_before = before;
        if (scene_ids) scene_ids[k] = id;
    }
    XDEM_REQUIRE(r.done(), "%s: trailing bytes", path);
    if (nscenes) *nscenes = ns_total;
    return XDEM_OK;
}

}  // extern "C"
