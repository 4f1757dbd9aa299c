// xdem_batch_kernels.hpp — device code of the fused FP32c path.
//
// One step of one scene (dem_simulation::oneStepAnalysis,
// source/simulation/dem_simulation.cpp:447-469, TCM = 1) becomes:
//   k_step   contacts (particle-particle in the reference traversal order, then
//            the scene's planes) + Verlet velocity update + next Verlet position
//            update + next cell key, written in the new sorted order (the
//            physical reorder rides on the writes);
//   k_patch / k_merge
//            re-sort by (cell key, particle id): only the particles whose cell
//            changed ("movers", recorded by k_step / k_pos together with their
//            insertion point) are sorted; the rest keep their order.  Up to
//            1024 movers, k_patch rewrites only the windows of positions they
//            shift (k_merge exits); otherwise k_patch ranks the movers and
//            clears the old cell ranges, and k_merge merges them with the rest
//            and writes the cell ranges (atomic minimum and maximum of the
//            positions per cell).  The result is the unique (key, id)
//            order, i.e. what the reference's stable sort of the id-ordered
//            keys produces (source/mphysics_cuda_def.cu:92-94).
#pragma once
#include "xdem_common.hpp"
#include "xdem_geom.hpp"

namespace xdem {
namespace fused {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;      // empty cell / sentinel key
constexpr int CMAX = 16;                     // contact-list slots per particle (LDS)
constexpr int HSLOTS = 16;                   // tangential-history slots per particle
constexpr int RTAB = 7;                      // neighbour rows kept in LDS during the walk
constexpr uint32_t WALL_KEY = 0x80000000u;   // history key of plane p: WALL_KEY | p
constexpr uint32_t DRUM_KEY = 0xC0000000u;   // drum q: DRUM_KEY | q << 2 | side
constexpr int LDS_T = 256;                   // lane stride of the per-thread LDS lists
constexpr int MERGE_LDS = 2048;              // sorted movers staged in LDS by k_merge up to this count
constexpr int ROWS_T = 256;                  // threads of the row kernels
constexpr int FEW = 1024;                    // movers re-placed by the window patch (k_patch)
constexpr int PATCH_T = 1024;                // threads of a k_patch block
constexpr int PATCH_GRID = 256;              // blocks of k_patch (grid-stride over the windows)
constexpr int RANK_MAX = 32768;              // movers sorted by counting ranks (k_patch, every block)
constexpr int RANK_CHUNK = 2048;             // mover keys staged in LDS per counting pass
constexpr int SWEEP_GRID = 1024;             // blocks of the grid-stride sort kernels
constexpr int VLK = 32;                      // candidate-list entries per particle
constexpr uint32_t VL_BAD = 0xFFFFFFFFu;     // no usable candidate list (walk every step)

// Scene constants (FP64 grid/time, FP32 material).
struct BScene {
    double wox, woy, woz, cs;
    double dt, half2dt;
    uint32_t gx, gy, gz, cell_base;
    uint32_t p0, n;           // segment in sorted order
    uint32_t plane0, nplane;  // into the plane table
    float gvx, gvy, gvz, dtf;
    float ckn, cvn, sr, mu;   // particle-particle: kn = ckn sqrt(R*), vn = sqrt(cvn M* kn)
    float mass_c;             // m = mass_c r^3 (rho 4/3 pi)
    int32_t rolling;
    uint32_t ncell;
    int32_t history;          // tangential-spring history on (DHS with a spring per contact)
    float cohpi;              // particle-particle cohesion x pi (0: none)
    double rmax;              // largest radius of the scene (cell culling)
    double pf_range;          // |coordinate - world origin| bound of the FP32 copy
    float pf_eps;             // distance error bound of the FP32 copy (2^-20 pf_range)
    float skin;               // candidate-list skin (distance beyond contact kept in a list)
    uint32_t drum0, ndrum;    // into the drum table (applied after the planes)
    uint32_t obj0, nobj;      // into the polygon-object and pose tables
    double inv_cs;            // 1 / cs (rounded): the cell of a coordinate by product first
};

struct BPlane {
    xdem_device_plane_info pl;
    float ckn, cvn, sr, mu;
};

// A rotating drum wall (xdem_drum_info) with its wall-pair constants.
struct BDrum {
    xdem_drum_info d;
    float ckn, cvn, sr, mu;
};

// A triangle-mesh object.  Its triangles are kept in the object's own frame,
// binned once by circumcentre into a dense local grid (cell size = the scene's,
// lo = grid corner), so a rigidly moving object is never re-binned: a particle
// is carried into the object frame and scans its 27 local cells.  With
// cs >= r_max + the largest circumradius every triangle within reach of a
// particle is in those cells, as in the reference's world grid.
struct BObject {
    double lox, loy, loz, cs;
    uint32_t gx, gy, gz, tri0;    // local grid; first triangle in the triangle table
    uint32_t cell0, ntri;         // first cell in the local cell table
    uint32_t step0;               // pose-step count when the scene was admitted
    uint32_t k_on, k_off;         // first and last step of the motion (the scene's steps,
                                  // from 1; k_off = 0xFFFFFFFF: no end)
    uint32_t pad;
    xdem_double3 origin0;         // pose at the scene's step 0
    xdem_double4 ep0;
    xdem_double3 vel, omega;      // the motion (zero: a static object)
    double dt;
    float ckn, cvn, sr, mu;
};

// Local-frame triangle and its face normal -(Q-P)x(R-P)/|.| (the contact
// normal of particle_polygonObject_collision_kernel, mphysics_cuda_impl.cuh:
// 923-1021).
struct BTri {
    double ccx, ccy, ccz, rc;     // circumcentre and circumradius (bounding sphere)
    double px, py, pz, qx, qy, qz, rx, ry, rz;
    float nx, ny, nz, pad;
};

// Pose of an object for the current step: origin, rows of A(ep), velocity.
struct BPose {
    double ox, oy, oz;
    double a[9];
    float vx, vy, vz, wx, wy, wz;
};

// One state buffer (SoA of packed records).  pr = (x, y, z, r) in FP64 (r is
// an FP32 value widened), vm = (v, m), wa = (omega, alpha.x), aa = (a, alpha.y),
// zi = (alpha.z bits, particle id), pf = the position relative to the scene's
// world origin in FP32 and r (x is NaN when a coordinate exceeds pf_range):
// 104 B per particle.  A neighbour candidate is screened on its 16-byte pf
// record; only candidates that pass load the 32-byte pr record.
struct BState {
    double4* pr;
    float4* vm;
    float4* wa;
    float4* aa;
    uint2* zi;
    float4* pf;
};

// The FP32 copy of a position.  In range, each coordinate is within
// 2^-24 pf_range of its FP64 value, so a distance computed from two copies is
// within 2^-21 pf_range of the true one; pf_eps = 2^-20 pf_range covers it
// twice.
__host__ __device__ inline float4 pf_of(const BScene& S, double x, double y, double z, float r) {
    const double rx = x - S.wox, ry = y - S.woy, rz = z - S.woz;
    const bool in = fabs(rx) <= S.pf_range && fabs(ry) <= S.pf_range && fabs(rz) <= S.pf_range;
    return make_float4(in ? (float)rx : __builtin_nanf(""), (float)ry, (float)rz, r);
}

// Element i of an array addressed with a 32-bit byte offset from its base
// (the SGPR base + VGPR offset form of the loads); the batch's capacities keep
// every per-particle and per-cell array below 4 GiB.
template <typename T>
__device__ inline const T& at32(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}

struct BlockEntry {
    uint32_t scene, j0, j1, pad;
};

// Movers (particles whose cell key changed) are recorded by the kernel that
// computes the new keys (k_step / k_pos), for the incremental sort: appended
// in arrival order with their insertion point in the previous sorted order,
// marked in moverbits and counted per ROWS_T block (the many-movers path of
// k_merge).  The marks and counts alternate between two buffers with the sort
// parity; each sort's k_patch clears the other parity's (the previous sort's)
// for the next one.
struct MoverOut {
    const uint32_t* key_old;   // sorted keys of the previous sort, storage order
    const uint2* crange;       // cell ranges of the previous sort
    uint32_t* mcount;
    uint64_t* mkey;            // (key << 32 | id)
    uint32_t* ms;              // storage index
    uint32_t* mplo;            // insertion point: previous entries below the new (key, id)
    uint4* mold;               // first FEW movers: (particles in the old cell, scene, old key,
                               // the old cell's first position)
    const uint2* zi;           // particle ids, storage order
    const uint32_t* perm;      // sorted position of the last sort -> storage index (null: identity)
    uint64_t* moverbits;
    uint32_t* blockcnt;
    uint32_t n, ncell_cap;
    uint32_t on;               // 0: keys only (a full sort follows)
};

struct StepArgs {
    BState in, out;
    const uint32_t* perm;      // sorted position -> storage index of `in`
    const uint2* crange;       // [key] = (first sorted position, one past last), EMPTY if none
    uint32_t* keynew;          // next cell key per sorted position
    MoverOut mv;               // movers of the next key (ADV)
    const BDrum* drums;
    const BObject* objs;
    const BTri* tris;
    const uint2* tcells;       // local cell table of the objects: (first, end) triangle
    const BPose* poses;
    const BScene* scenes;
    const BPlane* planes;
    const BlockEntry* blocks;
    uint32_t nblocks;
    uint32_t n;                // particles in the arrays (bounds of perm entries)
    // tangential-spring history (per storage index of `in`, per sorted
    // position of `out`): HSLOTS slots of (xi, key bits) and a count
    const float4* hin;
    const uint32_t* hcin;
    float4* hout;
    uint32_t* hcout;
    uint32_t* hoverflow;       // contacts found with every slot taken (not kept)
    // candidate lists (VLIST variants): per particle up to VLK entries in the
    // reference traversal order, lane-interleaved (entry k of storage index s
    // at vl[k * vl_stride + s]); entries name storage indices of `in` when
    // read and sorted positions (storage of `out`) when written
    uint32_t* vl;
    uint32_t* vc;              // entries per storage index; VL_BAD: none usable
    uint32_t vl_stride;
    const uint32_t* invperm;   // storage index of `in` -> sorted position
    const uint32_t* rebuild;   // per scene: rebuild every list this step
    const uint16_t* dirty;     // per cell key: tag of the last sort that saw a particle change
                               // cell within one cell of it
    const uint32_t* dtag;      // the last sort's tag
    uint32_t* wdisp;           // per block L of the launch: the largest displacement of this
                               // step (float bits, bounded above)
    uint32_t* force;           // per scene: set by the non-advancing variants that do not
                               // keep the lists (the next sort rebuilds every list)
    // parity probe outputs (per sorted position) — null in production
    float* dbg_f;              // 6 floats: force xyz, torque xyz
    uint32_t* dbg_cnt;
    uint32_t* dbg_part;        // dbg_max ids per particle
    uint32_t dbg_max;
};

__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nblk) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous run of logical blocks so neighbouring particle ranges share
    // one L2.
    uint32_t q = nblk / 8, r = nblk % 8, x = b % 8, i = b / 8;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// floor(fl(d / cs)), the reference's cell coordinate (calcGridPos), from the
// product with the rounded reciprocal when that is safely away from an
// integer: the product and the rounded quotient differ by at most 4 ulp of
// q (|q| < 2^26 here, so by less than 2^-24), so both have the same floor
// when the product's fraction lies in [2^-20, 1 - 2^-20]; otherwise (and for
// non-finite or large values) the division itself.
__device__ inline int cell_coord(double d, double cs, double inv_cs) {
    const double q = d * inv_cs;
    const double f = q - floor(q);
    if (f >= 0x1p-20 && f <= 1.0 - 0x1p-20 && fabs(q) < 0x1p26) return (int)floor(q);
    return (int)floor(d / cs);
}
__device__ inline void cell_of(const BScene& S, double x, double y, double z, int& cx, int& cy,
                               int& cz) {
    cx = cell_coord(x - S.wox, S.cs, S.inv_cs);
    cy = cell_coord(y - S.woy, S.cs, S.inv_cs);
    cz = cell_coord(z - S.woz, S.cs, S.inv_cs);
}

__device__ inline uint32_t key_of(const BScene& S, int cx, int cy, int cz) {
    uint32_t x = (uint32_t)cx & (S.gx - 1);
    uint32_t y = (uint32_t)cy & (S.gy - 1);
    uint32_t z = (uint32_t)cz & (S.gz - 1);
    return S.cell_base + (__umul24(__umul24(z, S.gy), S.gx) + __umul24(y, S.gx) + x);
}

// An upper bound of the largest u over the active lanes of the wave: the
// largest value of the high 16 bits any lane has (found bit by bit with
// ballots, so inactive lanes take no part), low bits set.  For non-negative
// floats' bit patterns the result read as a float is >= every lane's value
// (and a NaN, whose pattern exceeds every finite one, stays above them).
__device__ inline uint32_t wave_max_hi16(uint32_t u) {
    uint32_t cand = 0;
    for (int b = 31; b >= 16; --b) {
        const uint32_t t = cand | (1u << b);
        if (__ballot(u >= t)) cand = t;
    }
    return cand | 0xFFFFu;
}

__device__ inline uint64_t lanemask_lt64() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// First position of the previous sorted order whose key is >= kn, from the
// previous cell ranges (exact for the previous sort): the cell's own start,
// else the nearest non-empty cell within 8 keys, else a binary search.
__device__ inline uint32_t insertion_point(const uint2* crange, const uint32_t* key_old,
                                           uint32_t n, uint32_t ncell_cap, uint32_t kn) {
    if (kn < ncell_cap) {
        const uint2 c = crange[kn];
        if (c.x != EMPTY) return c.x;
        for (uint32_t d = 1; d <= 8; ++d) {
            if (kn >= d) {
                const uint2 a = crange[kn - d];
                if (a.x != EMPTY) return a.y;
            }
            if (kn + d < ncell_cap) {
                const uint2 b = crange[kn + d];
                if (b.x != EMPTY) return b.x;
            }
        }
    }
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key_old[mid] < kn)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Called by every active lane of a wave with its storage index, new key and
// particle id: one count reservation per wave that has movers.
__device__ inline void record_mover(const MoverOut& O, uint32_t s, uint32_t nw, uint32_t id,
                                    uint32_t scene) {
    const uint32_t ko = O.key_old[s];
    const bool mover = ko != nw;
    const uint64_t word = __ballot(mover);
    if (!word) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)word) - 1;
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(O.mcount, (uint32_t)__popcll(word));
    base = __shfl(base, (int)leader, 64);
    if (mover) {
        const uint32_t p = base + (uint32_t)__popcll(word & lanemask_lt64());
        if (p < O.n) {
            O.mkey[p] = ((uint64_t)nw << 32) | id;
            O.ms[p] = s;
            // past the entries of the new cell with a smaller id
            uint32_t lo = insertion_point(O.crange, O.key_old, O.n, O.ncell_cap, nw);
            while (lo < O.n && O.key_old[lo] == nw && O.zi[O.perm ? O.perm[lo] : lo].y < id) ++lo;
            O.mplo[p] = lo;
            if (p < (uint32_t)FEW) {
                uint32_t size = 0, first = 0;
                if (ko < O.ncell_cap) {
                    const uint2 c = O.crange[ko];
                    size = c.x == EMPTY ? 0u : c.y - c.x;
                    first = c.x == EMPTY ? 0u : c.x;
                }
                O.mold[p] = make_uint4(size, scene, ko, first);
            }
        }
        atomicOr((unsigned long long*)&O.moverbits[s >> 6], 1ull << (s & 63));
        atomicAdd(&O.blockcnt[s / ROWS_T], 1u);
    }
}

struct f3 {
    float x, y, z;
};
__device__ inline f3 F3(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ inline f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ inline void operator+=(f3& a, f3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}
__device__ inline float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 crossf(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// FP32 reciprocal and square root of the force law: the native instructions
// (v_rcp_f32, v_sqrt_f32; about 1 ulp) instead of the correctly rounded
// expansions — far inside the FP32c force tolerance, a fraction of the cost.
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// DHS force law in FP32 (mphysics_cuda_impl.cuh:379-401).  Ft and M keep the
// caller's values when the tangential velocity is exactly zero (as the
// reference kernel, whose Ft/M live outside the neighbour loop).  The
// reference's dot(dv, sh) is |e| exactly (sh = e/|e| and e is dv without its
// normal part); it is taken as |e| here: in FP32 the rounding residual of e
// along u, multiplied by |dv|/|e| in the dot product, would make the friction
// of a near head-on contact (|e| << |dv|) lose digits the reference keeps.
__device__ inline void dhs_f32(float kn, float vn, float ks, float vs, float mu, float dt,
                               float ir, float delta, f3 dv, f3 u, f3& Ft, f3& Fn, f3& M,
                               float fca = 0.0f) {
    float fsn = -kn * (delta * fsqrt(delta));
    float fdn = vn * dotf(dv, u);
    Fn = ((fsn + fca) + fdn) * u;
    f3 e = dv - dotf(dv, u) * u;
    float mag_e = fsqrt(dotf(e, e));
    if (mag_e != 0.0f) {
        float inv = frcp(mag_e);
        f3 sh = inv * e;
        float ds = mag_e * dt;
        float fn_len = fsqrt(dotf(Fn, Fn));
        Ft = fminf(ks * ds + vs * mag_e, mu * fn_len) * sh;
        M = crossf(ir * u, Ft);
    }
}

// DHS with a tangential spring (no reference counterpart; the reference's
// tangential force is memoryless).  xi is in/out: projected onto the current
// tangent plane, advanced by dt e, Ft = ks xi + vs e; when |Ft| > mu |Fn| the
// force is capped and the spring reset to (Ft - vs e) / ks.  With xi = 0 on
// entry this is dhs_f32's force whenever |e| != 0.
__device__ inline void hist_f32(float kn, float vn, float ks, float vs, float mu, float dt, float ir,
                                float delta, f3 dv, f3 u, f3& xi, f3& Ft, f3& Fn, f3& M,
                                float fca = 0.0f) {
    float fsn = -kn * (delta * fsqrt(delta));
    float fdn = vn * dotf(dv, u);
    Fn = ((fsn + fca) + fdn) * u;
    f3 e = dv - dotf(dv, u) * u;
    xi = xi - dotf(xi, u) * u;
    xi = xi + dt * e;
    Ft = ks * xi + vs * e;
    float ft = fsqrt(dotf(Ft, Ft)), fmax = mu * fsqrt(dotf(Fn, Fn));
    if (ft > fmax) {
        Ft = (fmax * frcp(ft)) * Ft;
        xi = ks > 0.0f ? frcp(ks) * (Ft - vs * e) : F3(0, 0, 0);
    }
    M = crossf(ir * u, Ft);
}

// One particle's history: the previous slots (keys staged in LDS, lane
// stride LDS_T) and the slots written this step, in evaluation order.
struct HistCtx {
    bool on;
    uint32_t* keys;
    const float4* in;
    float4* out;
    uint32_t* overflow;
    uint32_t nh, nout, hint;
    __device__ f3 lookup(uint32_t key) {
        // contacts come in nearly the same order every step: search from the
        // slot after the last match
        for (uint32_t q = 0; q < nh; ++q) {
            uint32_t k = hint + q;
            if (k >= nh) k -= nh;
            if (keys[k * LDS_T] == key) {
                hint = k + 1;
                const float4 v = in[k];
                return F3(v.x, v.y, v.z);
            }
        }
        return F3(0, 0, 0);
    }
    __device__ void store(uint32_t key, f3 xi) {
        if (nout < (uint32_t)HSLOTS) {
            out[nout] = make_float4(xi.x, xi.y, xi.z, __uint_as_float(key));
            ++nout;
        } else {
            atomicAdd(overflow, 1u);
        }
    }
};

struct Own {
    double x, y, z;
    f3 v, w;
    float r, m;
};

// Particle-particle contact of `o` with the particle stored at t of `in`
// (the predicate was already found true).  Accumulates into sumF/sumM.
template <bool HIST>
__device__ inline void pp_contact(const StepArgs& A, const BScene& S, const Own& o,
                                  const double4& jp, const float4& jvm, const float4& jwa,
                                  uint32_t jid, HistCtx& hc, f3& Ft, f3& Fn, f3& M, f3& sumF,
                                  f3& sumM) {
    double dx = jp.x - o.x, dy = jp.y - o.y, dz = jp.z - o.z;
    double d2 = dx * dx + dy * dy + dz * dz;
    // overlap (ri + rj) - |d| = ((ri + rj)^2 - |d|^2) / ((ri + rj) + |d|): the
    // cancelling difference in FP64, the rest in FP32 (no FP64 square root)
    double Sr = (double)o.r + jp.w;
    float dist = fsqrt((float)d2);
    float inv = frcp(dist);
    float cdist = (float)(Sr * Sr - d2) * frcp((float)Sr + dist);
    f3 u = F3((float)dx * inv, (float)dy * inv, (float)dz * inv);
    float jr = (float)jp.w;
    f3 jv = F3(jvm.x, jvm.y, jvm.z), jw = F3(jwa.x, jwa.y, jwa.z);
    f3 dv = S.rolling ? (jv + crossf(jw, (-jr) * u)) - (o.v + crossf(o.w, o.r * u)) : (jv - o.v);
    float R = (o.r * jr) * frcp(o.r + jr);
    float Ms = (o.m * jvm.w) * frcp(o.m + jvm.w);
    float kn = S.ckn * fsqrt(R);
    float vn = fsqrt(S.cvn * Ms * kn);
    // cohesionForce (mphysics_cuda_impl.cuh:259-279) is coh pi rc^2 with
    // rc = (3 R* |fsn| / (4 E*))^(1/3), the Hertz contact radius; with the DHS
    // stiffness kn = 4/3 E* sqrt(R*) and |fsn| = kn delta^1.5 that radius is
    // exactly sqrt(R* delta), so the term is coh pi R* delta (no cube root)
    const float fca = S.cohpi * R * cdist;
    if (HIST && hc.on) {
        f3 xi = hc.lookup(jid), cFt, cFn, cM;
        hist_f32(kn, vn, kn * S.sr, vn * S.sr, S.mu, S.dtf, o.r, cdist, dv, u, xi, cFt, cFn,
                 cM, fca);
        hc.store(jid, xi);
        sumF += cFn + cFt;
        sumM += cM;
        return;
    }
    dhs_f32(kn, vn, kn * S.sr, vn * S.sr, S.mu, S.dtf, o.r, cdist, dv, u, Ft, Fn, M, fca);
    sumF += Fn + Ft;
    sumM += M;
}

// The contacts of a list (lane stride LDS_T), in order.  PIPE: the records of
// contact c+1 are loaded before contact c is evaluated (more loads in flight,
// more registers).  HIST: the partner ids are loaded for the history keys.
template <bool PIPE, bool HIST>
__device__ inline void pp_contacts(const StepArgs& A, const BScene& S, const Own& o,
                                   const uint32_t* list, uint32_t cnt, HistCtx& hc, f3& Ft,
                                   f3& Fn, f3& M, f3& sumF, f3& sumM) {
    if (cnt == 0) return;
    uint32_t t = list[0];
    double4 jp = A.in.pr[t];
    float4 jvm = A.in.vm[t], jwa = A.in.wa[t];
    uint32_t jid = HIST ? A.in.zi[t].y : 0u;
    for (uint32_t c = 0; c < cnt; ++c) {
        if (PIPE) {
            const double4 cp = jp;
            const float4 cvm = jvm, cwa = jwa;
            const uint32_t cid = jid;
            if (c + 1 < cnt) {
                t = list[(c + 1) * LDS_T];
                jp = A.in.pr[t];
                jvm = A.in.vm[t];
                jwa = A.in.wa[t];
                if (HIST) jid = A.in.zi[t].y;
            }
            pp_contact<HIST>(A, S, o, cp, cvm, cwa, cid, hc, Ft, Fn, M, sumF, sumM);
        } else {
            if (c) {
                t = list[c * LDS_T];
                jp = A.in.pr[t];
                jvm = A.in.vm[t];
                jwa = A.in.wa[t];
                if (HIST) jid = A.in.zi[t].y;
            }
            pp_contact<HIST>(A, S, o, jp, jvm, jwa, jid, hc, Ft, Fn, M, sumF, sumM);
        }
    }
}

// UNR: candidates whose loads are issued together in the neighbour walk;
// PIPE: contact records loaded one contact ahead.  Both trade registers
// (waves per SIMD) for loads in flight per wave; the autotuner picks them
// together with the block size.  HIST: scenes with history on use the
// tangential-spring law (others the reference DHS law, same arithmetic as
// without HIST).
// Pose after tau seconds of an object's constant motion: origin0 + v tau,
// ep = q(omega, tau) * ep0 (Hamilton product, e0 first), A(ep) as makeTFM_1..3
// (mphysics_cuda_impl.cuh:624-638).
__host__ __device__ inline BPose pose_at(const BObject& ob, double tau) {
    BPose P;
    P.ox = ob.origin0.x + tau * ob.vel.x;
    P.oy = ob.origin0.y + tau * ob.vel.y;
    P.oz = ob.origin0.z + tau * ob.vel.z;
    const xdem_double3& w = ob.omega;
    const xdem_double4& p = ob.ep0;
    double e0 = p.x, e1 = p.y, e2 = p.z, e3 = p.w;
    const double wn = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    if (wn > 0) {
        const double h = 0.5 * (wn * tau);
        const double c = cos(h), sn = sin(h);
        const double q0 = c, q1 = sn * (w.x / wn), q2 = sn * (w.y / wn), q3 = sn * (w.z / wn);
        e0 = q0 * p.x - (q1 * p.y + q2 * p.z + q3 * p.w);
        e1 = q0 * p.y + p.x * q1 + (q2 * p.w - q3 * p.z);
        e2 = q0 * p.z + p.x * q2 + (q3 * p.y - q1 * p.w);
        e3 = q0 * p.w + p.x * q3 + (q1 * p.z - q2 * p.y);
    }
    P.a[0] = 2.0 * (e0 * e0 + e1 * e1 - 0.5);
    P.a[1] = 2.0 * (e1 * e2 - e0 * e3);
    P.a[2] = 2.0 * (e1 * e3 + e0 * e2);
    P.a[3] = 2.0 * (e1 * e2 + e0 * e3);
    P.a[4] = 2.0 * (e0 * e0 + e2 * e2 - 0.5);
    P.a[5] = 2.0 * (e2 * e3 - e0 * e1);
    P.a[6] = 2.0 * (e1 * e3 - e0 * e2);
    P.a[7] = 2.0 * (e2 * e3 + e0 * e1);
    P.a[8] = 2.0 * (e0 * e0 + e3 * e3 - 0.5);
    P.vx = (float)ob.vel.x;
    P.vy = (float)ob.vel.y;
    P.vz = (float)ob.vel.z;
    P.wx = (float)w.x;
    P.wy = (float)w.y;
    P.wz = (float)w.z;
    return P;
}

// Pose of an object at step k of its scene (k = 0: before the first step).
// The motion is active at steps k_on..k_off (the steps whose time k dt lies in
// [start_time, end_time]); the pose has moved by the active steps so far, and
// the object's velocity is the motion's while active and zero otherwise — the
// simulation object's per-step motion update (mo.n_act, xdem_host.hip).
__host__ __device__ inline BPose object_pose(const BObject& ob, uint32_t k) {
    const bool active = k >= ob.k_on && k <= ob.k_off;
    const uint32_t nact = k < ob.k_on ? 0u : (k < ob.k_off ? k : ob.k_off) - ob.k_on + 1u;
    BPose P = pose_at(ob, (double)nact * ob.dt);
    if (!active) P.vx = P.vy = P.vz = P.wx = P.wy = P.wz = 0.0f;
    return P;
}

// Poses of every object for the coming step (one block).  The step of a
// scene's objects is the number of pose updates since its admission, so the
// first step after admission is step 1, as the simulation object counts.
__global__ __launch_bounds__(256) void k_poses(const BObject* objs, BPose* poses, uint32_t nobj,
                                               uint32_t* counter) {
    const uint32_t c = *counter;
    for (uint32_t i = threadIdx.x; i < nobj; i += blockDim.x) {
        const BObject& ob = objs[i];
        poses[i] = object_pose(ob, c + 1 - ob.step0);
    }
    __syncthreads();
    if (threadIdx.x == 0) *counter = c + 1;
}

// Triangle contacts of one mesh object q of the scene with the particle `o`
// (Ft and M carried in and out across contacts, as in the step).  REACT: the
// reaction on the object is accumulated too: -(Fn + Ft) and its moment about
// the object origin, -(contact point - origin) x (Fn + Ft)
// (particle_polygonObject_collision_kernel, mphysics_cuda_impl.cuh:1008-1009).
template <bool REACT>
__device__ inline void mesh_object(const StepArgs& A, const BScene& S, const Own& o, double ri,
                                   uint32_t q, f3& Ft, f3& M, f3& F, f3* rf, f3* rm) {
    const d3 p = mk3(o.x, o.y, o.z);
    const BObject& ob = A.objs[S.obj0 + q];
    const BPose& P = A.poses[S.obj0 + q];
    const d3 dp = mk3(p.x - P.ox, p.y - P.oy, p.z - P.oz);
    // the particle in the object frame: A^T (p - origin)
    const d3 pl = mk3(P.a[0] * dp.x + P.a[3] * dp.y + P.a[6] * dp.z,
                      P.a[1] * dp.x + P.a[4] * dp.y + P.a[7] * dp.z,
                      P.a[2] * dp.x + P.a[5] * dp.y + P.a[8] * dp.z);
    // outside the local grid and its one-cell rim: no triangle in reach
    const double w = 2.0 * ob.cs;
    if (pl.x < ob.lox - w || pl.y < ob.loy - w || pl.z < ob.loz - w ||
        pl.x > ob.lox + ob.cs * ob.gx + w || pl.y > ob.loy + ob.cs * ob.gy + w ||
        pl.z > ob.loz + ob.cs * ob.gz + w)
        return;
    const int cx = (int)floor((pl.x - ob.lox) / ob.cs);
    const int cy = (int)floor((pl.y - ob.loy) / ob.cs);
    const int cz = (int)floor((pl.z - ob.loz) / ob.cs);
    if (cx < -1 || cy < -1 || cz < -1 || cx > (int)ob.gx || cy > (int)ob.gy || cz > (int)ob.gz)
        return;
    const float kn = ob.ckn * fsqrt(o.r);
    const float vn = fsqrt(ob.cvn * o.m * kn);
    const f3 vo = F3(P.vx, P.vy, P.vz), wo = F3(P.wx, P.wy, P.wz);
    for (int z = cz - 1; z <= cz + 1; ++z)
        for (int y = cy - 1; y <= cy + 1; ++y)
            for (int x = cx - 1; x <= cx + 1; ++x) {
                if (x < 0 || y < 0 || z < 0 || x >= (int)ob.gx || y >= (int)ob.gy ||
                    z >= (int)ob.gz)
                    continue;
                const uint2 cr = A.tcells[ob.cell0 + ((uint32_t)z * ob.gy + (uint32_t)y) * ob.gx +
                                          (uint32_t)x];
                const uint32_t e1 = min(cr.y, ob.ntri);
                for (uint32_t t = cr.x; t < e1; ++t) {
                    const BTri& T = A.tris[ob.tri0 + t];
                    // the triangle lies in its circumsphere: farther than rc + ri
                    // from the centre, no contact (1e-9 margin over the rounding)
                    const double sx = pl.x - T.ccx, sy = pl.y - T.ccy, sz = pl.z - T.ccz;
                    const double reach = T.rc + ri;
                    if (sx * sx + sy * sy + sz * sz > reach * reach * (1.0 + 1e-9)) continue;
                    const d3 cp = closest_pt_tri(pl, mk3(T.px, T.py, T.pz),
                                                 mk3(T.qx, T.qy, T.qz), mk3(T.rx, T.ry, T.rz));
                    const d3 dl = pl - cp;
                    const double cdist = ri - sqrt(dot(dl, dl));
                    if (!(cdist > 0)) continue;
                    // world normal A n and lever arm A cp (= contact point - origin)
                    const f3 u = F3((float)(P.a[0] * T.nx + P.a[1] * T.ny + P.a[2] * T.nz),
                                    (float)(P.a[3] * T.nx + P.a[4] * T.ny + P.a[5] * T.nz),
                                    (float)(P.a[6] * T.nx + P.a[7] * T.ny + P.a[8] * T.nz));
                    const f3 rc = F3((float)(P.a[0] * cp.x + P.a[1] * cp.y + P.a[2] * cp.z),
                                     (float)(P.a[3] * cp.x + P.a[4] * cp.y + P.a[5] * cp.z),
                                     (float)(P.a[6] * cp.x + P.a[7] * cp.y + P.a[8] * cp.z));
                    const f3 dv = S.rolling ? vo + crossf(wo, rc) - (o.v + crossf(o.w, o.r * u))
                                            : vo - o.v;
                    f3 Fn = F3(0, 0, 0);
                    dhs_f32(kn, vn, kn * ob.sr, vn * ob.sr, ob.mu, S.dtf, 0.0f, (float)cdist, dv,
                            u, Ft, Fn, M);
                    F += Fn + Ft;
                    if (REACT) {
                        *rf += -(Fn + Ft);
                        *rm += -crossf(rc, Fn + Ft);
                    }
                }
            }
}

// Drum walls (after the planes, in order) and triangle-mesh objects of the
// scene: FP64 geometry, FP32 DHS force with ir = 0 (no torque on the particle,
// as the reference's wall and triangle kernels).  Drum contacts are those of
// xdem_drum_contact_force_f64; a triangle contact is the reference's
// (distance to the closest point, face normal, object point velocity), with
// Ft carried from triangle contact to triangle contact as in
// particle_polygonObject_collision_kernel (mphysics_cuda_impl.cuh:923-1021).
// Drum wall q of the scene with the particle `o` (FP64 geometry, FP32 DHS with
// ir = 0).  REACT: the reaction on the drum is accumulated too: -(Fn + Ft) and
// its moment about the drum origin, (contact point - origin) x -(Fn + Ft), as
// xdem_drum_contact_force_f64 reports it.  HIST: each surface carries a
// tangential spring in the particle's history (key DRUM_KEY | q << 2 | side).
template <bool REACT, bool HIST = false>
__device__ inline void drum_wall(const StepArgs& A, const BScene& S, const Own& o, double ri,
                                 uint32_t q, f3& F, f3* rf, f3* rm, HistCtx* hc = nullptr) {
    const d3 p = mk3(o.x, o.y, o.z);
    const BDrum& bd = A.drums[S.drum0 + q];
    const d3 ax = of(bd.d.axis), c0 = of(bd.d.origin);
    const d3 dd = p - c0;
    const double a = dot(dd, ax);
    if (!(fabs(a) <= bd.d.half_length)) return;
    const d3 qv = dd - a * ax;
    const double q2 = dot(qv, qv);
    // clear of the shell (rho + r < R, decided on rho^2 with a margin far
    // above its rounding) and of the caps: no contact, no square root
    const double rin = bd.d.radius - ri;
    if (rin > 0 && q2 < rin * rin * (1.0 - 1e-12) &&
        !(bd.d.caps && ri - (bd.d.half_length - fabs(a)) > 0))
        return;
    const double rho = sqrt(q2);
    const float kn = bd.ckn * fsqrt(o.r);
    const float vn = fsqrt(bd.cvn * o.m * kn);
    // side: 0 shell, 1 the cap at +half_length, 2 the cap at -half_length
    auto contact = [&](d3 u, d3 cpt, double cdist, uint32_t side) {
        const f3 uf = F3((float)u.x, (float)u.y, (float)u.z);
        const d3 lever = cpt - c0;
        const d3 wv = bd.d.omega * cross(ax, lever);
        const f3 dv = F3((float)wv.x, (float)wv.y, (float)wv.z) - (o.v + crossf(o.w, o.r * uf));
        f3 Ft = F3(0, 0, 0), Fn = F3(0, 0, 0), M = F3(0, 0, 0);
        if constexpr (HIST) {
            const uint32_t key = DRUM_KEY | q << 2 | side;
            f3 xi = hc->lookup(key);
            hist_f32(kn, vn, kn * bd.sr, vn * bd.sr, bd.mu, S.dtf, 0.0f, (float)cdist, dv, uf, xi,
                     Ft, Fn, M);
            hc->store(key, xi);
        } else {
            dhs_f32(kn, vn, kn * bd.sr, vn * bd.sr, bd.mu, S.dtf, 0.0f, (float)cdist, dv, uf, Ft,
                    Fn, M);
        }
        F += Fn + Ft;
        if (REACT) {
            *rf += -(Fn + Ft);
            *rm += crossf(F3((float)lever.x, (float)lever.y, (float)lever.z), -(Fn + Ft));
        }
    };
    const double cs = ri + rho - bd.d.radius;
    if (cs > 0 && rho > 0) {
        const d3 u = qv / rho;
        contact(u, c0 + a * ax + bd.d.radius * u, cs, 0u);
    }
    const double cc = ri - (bd.d.half_length - fabs(a));
    if (bd.d.caps && cc > 0 && rho < bd.d.radius) {
        const double sg = a >= 0 ? 1.0 : -1.0;
        contact(sg * ax, c0 + (sg * bd.d.half_length) * ax + qv, cc, a >= 0 ? 1u : 2u);
    }
}

__device__ inline void walls_and_meshes(const StepArgs& A, const BScene& S, const Own& o,
                                        double ri, f3& F) {
    for (uint32_t q = 0; q < S.ndrum; ++q) drum_wall<false>(A, S, o, ri, q, F, nullptr, nullptr);
    f3 Ft = F3(0, 0, 0), M = F3(0, 0, 0);  // carried across triangle contacts
    for (uint32_t q = 0; q < S.nobj; ++q) mesh_object<false>(A, S, o, ri, q, Ft, M, F, nullptr, nullptr);
}

// Waves per SIMD the register allocation targets for each variant (the
// counts its register footprint nearly reaches without spilling).
constexpr int step_waves(int unr, bool pipe, bool hist, bool geom) {
    return hist && geom ? 3 : hist || pipe || geom ? 4 : 5;
}

// GEOM: the variant that also evaluates drum walls and triangle meshes (used
// when some scene of the batch has them; the others carry no registers for it).
template <bool ADV, bool DBG, int UNR, bool PIPE, bool HIST, bool GEOM = false, bool FLAT = true,
          bool VLIST = false>
__global__ __launch_bounds__(256, step_waves(UNR, PIPE, HIST, GEOM)) void k_step(StepArgs A) {
    static_assert(!VLIST || (!DBG && !HIST && !GEOM && FLAT),
                  "candidate lists: plane-only per-lane-walk variants");
    // per-thread LDS lists, lane-contiguous (entry c of thread t at c * LDS_T + t)
    __shared__ uint32_t clist[LDS_T * CMAX];
    __shared__ uint32_t hkeys[HIST ? LDS_T * HSLOTS : 1];
    // the non-empty neighbour rows after the first two (per thread, lane stride
    // LDS_T); the history variants borrow the thread's own key slots (row i in
    // slots 2i and 2i+1), which the thread fills only after its walk, so no
    // thread ever touches another thread's slots
    static_assert(2 * RTAB <= HSLOTS, "row table must fit in the history keys");
    __shared__ uint2 rtab_own[HIST || !FLAT ? 1 : LDS_T * RTAB];
    // candidate lists: the block's largest displacement and the lanes done
    __shared__ uint32_t sdisp[2];
    if constexpr (VLIST) {
        if (threadIdx.x == 0) {
            sdisp[0] = 0;
            sdisp[1] = 0;
        }
        __syncthreads();
    }
    auto rtab_put = [&](uint32_t i, uint32_t rb, uint32_t re) {
        if constexpr (HIST) {
            hkeys[(2 * i) * LDS_T + threadIdx.x] = rb;
            hkeys[(2 * i + 1) * LDS_T + threadIdx.x] = re;
        } else {
            rtab_own[i * LDS_T + threadIdx.x] = make_uint2(rb, re);
        }
    };
    auto rtab_get = [&](uint32_t i) -> uint2 {
        if constexpr (HIST)
            return make_uint2(hkeys[(2 * i) * LDS_T + threadIdx.x],
                              hkeys[(2 * i + 1) * LDS_T + threadIdx.x]);
        else
            return rtab_own[i * LDS_T + threadIdx.x];
    };
    const uint32_t L = xcd_remap(blockIdx.x, gridDim.x);
    const BlockEntry be = A.blocks[L];
    const BScene S = A.scenes[be.scene];
    const uint32_t j = be.j0 + threadIdx.x;
    // a variant without lists leaves them stale: the next sort rebuilds them.
    // Every step block ends with a step that does not advance, so marking in
    // those variants covers a whole block (and keeps the advancing variants'
    // registers as they are)
    if constexpr (!VLIST && !ADV)
        if (A.force && threadIdx.x == 0) A.force[be.scene] = 1u;
    // Every lane arrives once with its wave's largest displacement (0 for a
    // lane without a particle); the lane that completes the block's count
    // writes the block's maximum (LDS atomics, one per wave and path).
    auto disp_arrive = [&](uint32_t m) {
        const uint64_t act = __ballot(1);
        if (__lane_id() == (uint32_t)__ffsll((unsigned long long)act) - 1) {
            if (m) atomicMax(&sdisp[0], m);
            const uint32_t k = (uint32_t)__popcll(act);
            if (atomicAdd(&sdisp[1], k) + k == blockDim.x) A.wdisp[L] = atomicMax(&sdisp[0], 0u);
        }
    };
    if (j >= be.j1 || j >= A.n) {
        if constexpr (VLIST) disp_arrive(0u);
        return;
    }
    uint32_t* mylist = clist + threadIdx.x;
    const uint32_t s = A.perm[j];
    if (s >= A.n) {  // defensive: perm is a permutation of [0, n)
        if constexpr (VLIST) disp_arrive(0u);
        return;
    }
    const double4 ip = A.in.pr[s];
    // The 5-wave variants load the rest of the own record after the neighbour
    // walk (the walk does not hold it in registers); the 4-wave ones have the
    // registers and load it here, one latency less.
    constexpr bool EARLY = step_waves(UNR, PIPE, HIST, GEOM) < 5;
    float4 vm = make_float4(0, 0, 0, 0), wa = vm, aa = vm;
    uint2 zi = make_uint2(0, 0);
    if (EARLY) {
        vm = A.in.vm[s];
        wa = A.in.wa[s];
        aa = A.in.aa[s];
        zi = A.in.zi[s];
    }
    HistCtx hc{};
    if (HIST) {
        hc.on = S.history != 0;
        hc.keys = hkeys + threadIdx.x;
        hc.in = A.hin + (size_t)s * HSLOTS;
        hc.out = A.hout + (size_t)j * HSLOTS;
        hc.overflow = A.hoverflow;
        hc.nh = hc.on ? min(A.hcin[s], (uint32_t)HSLOTS) : 0u;
    }
    Own o;
    o.x = ip.x;
    o.y = ip.y;
    o.z = ip.z;
    o.r = (float)ip.w;
    int cx, cy, cz;
    cell_of(S, o.x, o.y, o.z, cx, cy, cz);
    // Candidate lists (VLIST): a lane reuses its list (the candidates within
    // ri + rj + skin when it was built, in the reference traversal order)
    // unless the scene rebuilds every list this step (displacement bound past
    // half the skin, or a block start), the particle moved in the sorted
    // order (its list slot moves with it), a particle changed cell within one
    // cell of its own (the traversal order around it may have changed, or a
    // new candidate arrived), or it has no usable list.  Otherwise its
    // contacts and their order are those of the walk.
    bool rebuild = true;
    uint32_t vcnt = 0;
    if constexpr (VLIST) {
        vcnt = A.vc[s];
        const uint32_t kown = A.mv.key_old[j];
        const bool dirty = A.dirty[kown] == *A.dtag;
        rebuild = A.rebuild[be.scene] != 0 || s != j || dirty || vcnt > (uint32_t)VLK;
    }
    f3 Ft = F3(0, 0, 0), Fn = F3(0, 0, 0), M = F3(0, 0, 0);
    f3 sumF = F3(0, 0, 0), sumM = F3(0, 0, 0);
    uint32_t ntotal = 0;
    const double ri = (double)o.r;
    // Cell culling: a particle j touching i has |pj - pi| < ri + rj <= ri +
    // rmax <= cell size, so it lies in a real-space neighbour cell of i, and
    // every point of that cell is at least the face distances below away from
    // pi.  A cell whose squared distance bound reaches (ri + rmax)^2 (1e-4
    // margin, far above the FP32 rounding of the offsets and the rounding of
    // the cell assignment) holds no contact of i and is not visited; skipping
    // it changes neither the contacts nor their order.
    const float csf = (float)S.cs;
    // a list build also keeps the candidates within the skin
    const double reach = ri + S.rmax + (VLIST ? (double)S.skin : 0.0);
    const float reach2 = (float)(reach * reach) * (1.0f + 1e-4f);
    const float lox = (float)(o.x - (S.wox + (double)cx * S.cs));
    const float loy = (float)(o.y - (S.woy + (double)cy * S.cs));
    const float loz = (float)(o.z - (S.woz + (double)cz * S.cs));
    auto face2 = [&](int d, float lo) {
        const float b = d < 0 ? lo : d > 0 ? csf - lo : 0.0f;
        return b * b;
    };
    // FP32 screen of a candidate: rejects only when the FP32 distance exceeds
    // ri + rj by more than twice its error bound (then the FP64 predicate is
    // false); out-of-range copies (x NaN, so d2 is NaN) are never rejected.
    const float4 ownf = pf_of(S, o.x, o.y, o.z, o.r);
    const float sr0 = ownf.w + S.pf_eps;
    auto screen_d2 = [&](const float4& f) {
        const float dx = f.x - ownf.x, dy = f.y - ownf.y, dz = f.z - ownf.z;
        return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    };
    auto may_touch = [&](const float4& f) {
        const float d2 = screen_d2(f);
        const float sr = sr0 + f.w;
        return !(d2 >= sr * sr * (1.0f + 0x1p-19f));
    };
    // within ri + rj + skin (list builds; same widened FP32 bound)
    auto in_skin = [&](const float4& f) {
        const float d2 = screen_d2(f);
        const float sr = sr0 + f.w + S.skin;
        return !(d2 >= sr * sr * (1.0f + 0x1p-19f));
    };
    // phase 1: candidates in the reference order (z, y, x cells; ascending j).
    // The first CMAX that pass the FP64 predicate go to this thread's LDS list
    // (phase 2 evaluates them in order); the rest are only counted, and a
    // particle with more contacts rescans its cells for each further chunk of
    // CMAX (rare).  Ft and M carry over from contact to contact in list order,
    // so the chunking does not change the result.  Contact evaluation stays out
    // of the candidate loop, which keeps the loop's register footprint small.
    uint32_t skip = 0;
    auto consider = [&](uint32_t t, const double4& jp) {
        double ex = jp.x - o.x, ey = jp.y - o.y, ez = jp.z - o.z;
        double d2 = ex * ex + ey * ey + ez * ez;
        double Sr = ri + jp.w, S2 = Sr * Sr;
        if (d2 > S2 * 1.000001) return;
        // the reference predicate (Sr - fl(sqrt(d2)) > 0) is decided by d2
        // alone unless d2 is within 1e-9 of Sr^2 (fl(sqrt) is within half an
        // ulp of the root); only then is the square root taken
        if (!(d2 < S2 * (1.0 - 1e-9)) && !(Sr - sqrt(d2) > 0)) return;
        if (DBG && skip == 0 && ntotal < A.dbg_max)
            A.dbg_part[(size_t)j * A.dbg_max + ntotal] = A.in.zi[t].y;
        if (ntotal >= skip && ntotal - skip < (uint32_t)CMAX) mylist[(ntotal - skip) * LDS_T] = t;
        ++ntotal;
    };
    // the 27 cells one by one (x wrap, and the rescans)
    auto scan_cells = [&]() {
#pragma unroll 1
        for (int dz = -1; dz <= 1; ++dz)
#pragma unroll 1
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll 1
                for (int dx = -1; dx <= 1; ++dx) {
                    if (face2(dx, lox) + face2(dy, loy) + face2(dz, loz) >= reach2) continue;
                    const uint2 cr = at32(A.crange, key_of(S, cx + dx, cy + dy, cz + dz));
                    if (cr.x == EMPTY) continue;
                    const uint32_t e1 = min(cr.y, A.n);
                    for (uint32_t jj = cr.x; jj < e1; ++jj) {
                        if (jj == j) continue;
                        const uint32_t t = at32(A.perm, jj);
                        if (t >= A.n) continue;
                        if (may_touch(at32(A.in.pf, t))) consider(t, at32(A.in.pr, t));
                    }
                }
    };
    if (cx >= 1 && cx + 1 < (int)S.gx) {
        // Away from the x wrap the three cells of an x-row have consecutive keys,
        // so their particles are one contiguous run of sorted positions.  A row
        // is skipped when even its middle cell is out of reach; its end cells
        // are dropped when out of reach.  Both walks visit the rows in order and
        // each row in ascending sorted position (the order of the cell loop) and
        // load UNR index words, then UNR FP32 screening records, per chunk; the
        // FP64 records of the candidates that pass the screen are loaded for the
        // exact test.
        //   FLAT: the non-empty rows (at most 9) are listed first, the first two
        // in registers and the rest in LDS, and each lane walks its rows as one
        // stream: a slot past the end of the current row takes the next row's
        // candidates, so a chunk may cross one row boundary and lanes with short
        // rows do not wait for the longest row of the wave.
        //   Row walk: the wave walks row r (the range of row r+1 loaded ahead)
        // until its longest lane is done, then row r+1.
        auto row_range = [&](int r) {
            const int dy = (r % 3) - 1, dz = (r / 3) - 1;
            const float b2 = face2(dy, loy) + face2(dz, loz);
            if (b2 >= reach2) return make_uint2(EMPTY, 0);
            const bool m1 = face2(-1, lox) + b2 < reach2, p1 = face2(1, lox) + b2 < reach2;
            const uint32_t k1 = key_of(S, cx, cy + dy, cz + dz);
            const uint2 c0 = m1 ? at32(A.crange, k1 - 1) : make_uint2(EMPTY, 0);
            const uint2 c1 = at32(A.crange, k1);
            const uint2 c2 = p1 ? at32(A.crange, k1 + 1) : make_uint2(EMPTY, 0);
            uint2 rg;
            rg.x = c0.x != EMPTY ? c0.x : (c1.x != EMPTY ? c1.x : c2.x);
            rg.y = c2.x != EMPTY ? c2.y : (c1.x != EMPTY ? c1.y : c0.y);
            return rg;
        };
        if constexpr (FLAT) {
            // jj..e1: the current row's remaining run; nb..ne: the next row's run
            uint32_t jj = 0, e1 = 0, nb = 0, ne = 0, nt = 0;
            uint32_t lcnt = 0;  // list entries written (VLIST rebuild)
            // the 27 range loads are issued unconditionally (key_of wraps into
            // the scene's grid and 1 <= cx <= gx - 2, so every key is in range)
            // and the culled cells are masked afterwards
            if (!VLIST || rebuild)
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const int dy = (r % 3) - 1, dz = (r / 3) - 1;
                const uint32_t k1 = key_of(S, cx, cy + dy, cz + dz);
                const uint2 c0 = at32(A.crange, k1 - 1), c1 = at32(A.crange, k1),
                            c2 = at32(A.crange, k1 + 1);
                const float b2 = face2(dy, loy) + face2(dz, loz);
                const bool alive = b2 < reach2;
                const bool h0 = alive && face2(-1, lox) + b2 < reach2 && c0.x != EMPTY;
                const bool h1 = alive && c1.x != EMPTY;
                const bool h2 = alive && face2(1, lox) + b2 < reach2 && c2.x != EMPTY;
                const uint32_t rb = h0 ? c0.x : (h1 ? c1.x : (h2 ? c2.x : EMPTY));
                const uint32_t re = min(h2 ? c2.y : (h1 ? c1.y : c0.y), A.n);
                if (rb == EMPTY || rb >= re) continue;
                jj = nt == 0 ? rb : jj;
                e1 = nt == 0 ? re : e1;
                nb = nt == 1 ? rb : nb;
                ne = nt == 1 ? re : ne;
                if (nt >= 2) rtab_put(nt - 2, rb, re);
                ++nt;
            }
            const uint32_t ntab = nt > 2 ? nt - 2 : 0;
            uint32_t ti = 0;  // next row table entry
#pragma unroll 1
            while (jj < e1) {
                uint32_t t[UNR];
                uint32_t qq[UNR];
                float4 f[UNR];
                const uint32_t nbase = nb - e1;  // slot k >= e1 maps to nb + (k - e1)
                // idle slots (past both rows, or the thread itself) load the
                // thread's own entries, so the loads need no branch; ok[]
                // keeps them out of the screen
                bool ok[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const uint32_t k = jj + u;
                    const uint32_t q = k < e1 ? k : nbase + k;
                    ok[u] = (k < e1 || q < ne) && q != j;  // ne <= n
                    qq[u] = q;
                    t[u] = at32(A.perm, ok[u] ? q : j);
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    ok[u] = ok[u] && t[u] < A.n;
                    f[u] = at32(A.in.pf, min(t[u], A.n - 1));
                }
                // advance: stay in the current row, continue inside the next one,
                // or (that one used up too) start the next chunk at the row after it
                jj += UNR;
                if (jj >= e1) {
                    const uint32_t c2 = jj - e1;
                    if (c2 < ne - nb) {
                        jj = nb + c2;
                        e1 = ne;
                        nb = ne = 0;
                        if (ti < ntab) {
                            const uint2 rg = rtab_get(ti++);
                            nb = rg.x;
                            ne = rg.y;
                        }
                    } else {
                        jj = e1 = nb = ne = 0;
                        if (ti < ntab) {
                            const uint2 rg = rtab_get(ti++);
                            jj = rg.x;
                            e1 = rg.y;
                        }
                        if (ti < ntab) {
                            const uint2 rg = rtab_get(ti++);
                            nb = rg.x;
                            ne = rg.y;
                        }
                    }
                }
                if constexpr (VLIST) {
                    // the list keeps sorted positions: the storage indices of `out`
#pragma unroll
                    for (int u = 0; u < UNR; ++u)
                        if (ok[u] && in_skin(f[u])) {
                            if (lcnt < (uint32_t)VLK) A.vl[(size_t)lcnt * A.vl_stride + j] = qq[u];
                            ++lcnt;
                        }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u)
                    if (ok[u] && may_touch(f[u])) consider(t[u], at32(A.in.pr, t[u]));
            }
            if constexpr (VLIST) {
                if (rebuild) {
                    A.vc[j] = lcnt > (uint32_t)VLK ? VL_BAD : lcnt;
                } else {
                    // the list: entries are storage indices of `in`; each is
                    // rewritten with its sorted position (its storage index in
                    // `out`) where that differs, in place (slot s == j)
                    // LU entries per batch; the next batch's entries are loaded
                    // while this batch's records are in flight
                    constexpr int LU = UNR < 2 ? 2 : UNR;
                    uint32_t en[LU];
#pragma unroll
                    for (int u = 0; u < LU; ++u)
                        en[u] = (uint32_t)u < vcnt ? A.vl[(size_t)u * A.vl_stride + s] : s;
#pragma unroll 1
                    for (uint32_t k = 0; k < vcnt; k += LU) {
                        uint32_t e[LU], nx2[LU];
                        float4 f[LU];
#pragma unroll
                        for (int u = 0; u < LU; ++u) {
                            e[u] = min(en[u], A.n - 1);  // defensive: entries are < n
                            f[u] = at32(A.in.pf, e[u]);
                            nx2[u] = at32(A.invperm, e[u]);
                        }
#pragma unroll
                        for (int u = 0; u < LU; ++u) {
                            const uint32_t kn = k + LU + u;
                            en[u] = kn < vcnt ? A.vl[(size_t)kn * A.vl_stride + s] : s;
                        }
#pragma unroll
                        for (int u = 0; u < LU; ++u)
                            if (k + u < vcnt) {
                                if (nx2[u] != e[u])
                                    A.vl[(size_t)(k + u) * A.vl_stride + s] = nx2[u];
                                if (may_touch(f[u])) consider(e[u], at32(A.in.pr, e[u]));
                            }
                    }
                }
            }
        } else {
            uint2 next = row_range(0);
#pragma unroll 1
            for (int r = 0; r < 9; ++r) {
                const uint2 row = next;
                if (r < 8) next = row_range(r + 1);
                if (row.x == EMPTY) continue;
                const uint32_t e1 = min(row.y, A.n);
                for (uint32_t jj = row.x; jj < e1; jj += UNR) {
                    uint32_t t[UNR];
                    float4 f[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        const uint32_t k = jj + u;
                        const bool ok = k < e1 && k != j;  // e1 <= n
                        t[u] = ok ? at32(A.perm, k) : A.n;
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u)
                        if (t[u] < A.n) f[u] = at32(A.in.pf, t[u]);
#pragma unroll
                    for (int u = 0; u < UNR; ++u)
                        if (t[u] < A.n && may_touch(f[u])) consider(t[u], at32(A.in.pr, t[u]));
                }
            }
        }
    } else {
        scan_cells();
        if constexpr (VLIST) A.vc[j] = VL_BAD;  // next to the x wrap: no list, always walked
    }
    if (!EARLY) {
        vm = A.in.vm[s];
        wa = A.in.wa[s];
    }
    o.v = F3(vm.x, vm.y, vm.z);
    o.w = F3(wa.x, wa.y, wa.z);
    o.m = vm.w;
    if (HIST)
        for (uint32_t k = 0; k < hc.nh; ++k) hc.keys[k * LDS_T] = __float_as_uint(hc.in[k].w);
    // phase 2: the listed contacts, in order; then the further chunks
    const uint32_t found = ntotal;
    pp_contacts<PIPE, HIST>(A, S, o, mylist, min(found, (uint32_t)CMAX), hc, Ft, Fn, M, sumF, sumM);
#pragma unroll 1
    for (skip = CMAX; skip < found; skip += CMAX) {
        ntotal = 0;
        scan_cells();
        pp_contacts<PIPE, HIST>(A, S, o, mylist, min(found - skip, (uint32_t)CMAX), hc, Ft, Fn, M,
                                sumF, sumM);
    }
    ntotal = found;
    // F = m g + particle-particle (vv_update_velocity reset, then cu_calculate_p2p)
    f3 F = F3(o.m * S.gvx, o.m * S.gvy, o.m * S.gvz) + sumF;
    f3 T = sumM;
    // planes, in order (plane_contact_force_kernel: no torque, ir = 0)
    for (uint32_t p = 0; p < S.nplane; ++p) {
        const BPlane& bp = A.planes[S.plane0 + p];
        d3 xp = mk3(o.x, o.y, o.z);
        d3 dp = xp - of(bp.pl.xw);
        d3 unit = mk3(0, 0, 0);
        d3 wp = mk3(dot(dp, of(bp.pl.u1)), dot(dp, of(bp.pl.u2)), dot(dp, of(bp.pl.uw)));
        double cdist = plane_detect<false>(bp.pl, xp, wp, unit, ri);
        if (cdist > 0) {
            f3 u = F3((float)unit.x, (float)unit.y, (float)unit.z);
            f3 dv = S.rolling ? -(o.v + crossf(o.w, o.r * u)) : -o.v;
            float kn = bp.ckn * fsqrt(o.r);
            float vn = fsqrt(bp.cvn * o.m * kn);
            f3 pFt = F3(0, 0, 0), pFn = F3(0, 0, 0), pM = F3(0, 0, 0);
            if (HIST && hc.on) {
                f3 xi = hc.lookup(WALL_KEY | p);
                hist_f32(kn, vn, kn * bp.sr, vn * bp.sr, bp.mu, S.dtf, 0.0f, (float)cdist, dv, u,
                         xi, pFt, pFn, pM);
                hc.store(WALL_KEY | p, xi);
            } else {
                dhs_f32(kn, vn, kn * bp.sr, vn * bp.sr, bp.mu, S.dtf, 0.0f, (float)cdist, dv, u,
                        pFt, pFn, pM);
            }
            F += pFn + pFt;
        }
    }
    if (GEOM && (S.ndrum | S.nobj)) {
        if (HIST && hc.on) {
            // a scene with history has drums and no meshes: drum springs
            // after the plane springs
            for (uint32_t q = 0; q < S.ndrum; ++q)
                drum_wall<false, HIST>(A, S, o, ri, q, F, nullptr, nullptr, &hc);
        } else {
            walls_and_meshes(A, S, o, ri, F);
        }
    }
    if (DBG) {
        float* d = A.dbg_f + 6 * (size_t)j;
        d[0] = F.x - o.m * S.gvx;
        d[1] = F.y - o.m * S.gvy;
        d[2] = F.z - o.m * S.gvz;
        d[3] = T.x;
        d[4] = T.y;
        d[5] = T.z;
        A.dbg_cnt[j] = ntotal;
    }
    // vv_update_velocity_kernel (mphysics_cuda_impl.cuh:98-134) in FP32
    if (!EARLY) {
        aa = A.in.aa[s];
        zi = A.in.zi[s];
    }
    float inv_m = frcp(o.m);
    float inv_i = frcp(0.4f * o.m * o.r * o.r);
    f3 a0 = F3(aa.x, aa.y, aa.z), al0 = F3(wa.w, aa.w, __uint_as_float(zi.x));
    f3 a = inv_m * F, al = inv_i * T;
    float hdt = 0.5f * S.dtf;
    f3 v = o.v + hdt * (a0 + a);
    f3 w = o.w + hdt * (al0 + al);
    double nx = o.x, ny = o.y, nz = o.z;
    if (ADV) {
        // next vv_update_position (:85-96) in FP64: p += dt v + half2dt a
        nx += S.dt * (double)v.x + S.half2dt * (double)a.x;
        ny += S.dt * (double)v.y + S.half2dt * (double)a.y;
        nz += S.dt * (double)v.z + S.half2dt * (double)a.z;
        int kx, ky, kz;
        cell_of(S, nx, ny, nz, kx, ky, kz);
        const uint32_t kn = key_of(S, kx, ky, kz);
        A.keynew[j] = kn;
        if (A.mv.on) record_mover(A.mv, j, kn, zi.y, be.scene);
        if constexpr (VLIST) {
            // the scene's largest displacement of the step (bounded above)
            const float ddx = (float)(nx - o.x), ddy = (float)(ny - o.y), ddz = (float)(nz - o.z);
            const float d = fsqrt(ddx * ddx + ddy * ddy + ddz * ddz) * 1.0001f;
            disp_arrive(wave_max_hi16(__float_as_uint(d)));
        }
    } else if constexpr (VLIST) {
        disp_arrive(0u);  // no advance (a block's last step)
    }
    A.out.pr[j] = make_double4(nx, ny, nz, ip.w);
    A.out.pf[j] = pf_of(S, nx, ny, nz, o.r);
    A.out.vm[j] = make_float4(v.x, v.y, v.z, o.m);
    A.out.wa[j] = make_float4(w.x, w.y, w.z, al.x);
    A.out.aa[j] = make_float4(a.x, a.y, a.z, al.y);
    A.out.zi[j] = make_uint2(__float_as_uint(al.z), zi.y);
    if (HIST) A.hcout[j] = hc.nout;
}

// Reactions of the mesh objects and drums for one step: their contacts of
// k_step re-evaluated on the same state and poses, in the same order (Ft and M
// carried across the objects of a particle), so the forces are the step's own.
// Per body, the block's particles are reduced in a fixed order in LDS (FP64)
// into part[(L * slots + q) * 6 ...]; every lane stays to the end.
__global__ __launch_bounds__(256) void k_react(StepArgs A, double* part, uint32_t slots) {
    __shared__ double red[6][256];
    const uint32_t L = xcd_remap(blockIdx.x, gridDim.x);
    const BlockEntry be = A.blocks[L];
    const BScene S = A.scenes[be.scene];
    const uint32_t t = threadIdx.x, j = be.j0 + t;
    bool live = j < be.j1 && j < A.n;
    const uint32_t s = live ? A.perm[j] : 0u;
    live = live && s < A.n;
    Own o{};
    if (live) {
        const double4 ip = A.in.pr[s];
        const float4 vm = A.in.vm[s], wa = A.in.wa[s];
        o.x = ip.x;
        o.y = ip.y;
        o.z = ip.z;
        o.r = (float)ip.w;
        o.v = F3(vm.x, vm.y, vm.z);
        o.w = F3(wa.x, wa.y, wa.z);
        o.m = vm.w;
    }
    const double ri = (double)o.r;
    f3 Ft = F3(0, 0, 0), M = F3(0, 0, 0), F = F3(0, 0, 0);
    // bodies of the scene: its mesh objects, then its drums (slot nobj + d)
    for (uint32_t q = 0; q < S.nobj + S.ndrum && q < slots; ++q) {  // uniform in the block
        f3 rf = F3(0, 0, 0), rm = F3(0, 0, 0);
        if (live) {
            if (q < S.nobj)
                mesh_object<true>(A, S, o, ri, q, Ft, M, F, &rf, &rm);
            else
                drum_wall<true>(A, S, o, ri, q - S.nobj, F, &rf, &rm);
        }
        red[0][t] = rf.x;
        red[1][t] = rf.y;
        red[2][t] = rf.z;
        red[3][t] = rm.x;
        red[4][t] = rm.y;
        red[5][t] = rm.z;
        __syncthreads();
        for (uint32_t h = blockDim.x >> 1; h > 0; h >>= 1) {
            if (t < h)
                for (int k = 0; k < 6; ++k) red[k][t] += red[k][t + h];
            __syncthreads();
        }
        if (t < 6) part[((size_t)L * slots + q) * 6 + t] = red[t][0];
        __syncthreads();
    }
}

// Per object g (robj[g] = first and end block of its scene, slot q): the block
// partials summed in a fixed order into out[6 g ...] (force, moment).
__global__ __launch_bounds__(256) void k_react_sum(const double* part, uint32_t slots,
                                                   const uint4* robj, double* out) {
    __shared__ double red[6][256];
    const uint4 r = robj[blockIdx.x];
    const uint32_t t = threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t L = r.x + t; L < r.y; L += 256)
        for (int k = 0; k < 6; ++k) acc[k] += part[((size_t)L * slots + r.z) * 6 + k];
    for (int k = 0; k < 6; ++k) red[k][t] = acc[k];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 6; ++k) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t < 6) out[(size_t)blockIdx.x * 6 + t] = red[t][0];
}

// Verlet position update (optional) + cell key of every particle, in place,
// in storage order (block start / parity prepare).  force: every candidate
// list of the block's scene is rebuilt at the next sort.  wdisp (with ADV,
// when the block table is the contact kernel's): the block's largest
// displacement instead, as the contact kernel records it, so that the lists
// carry over the block start.
template <bool ADV>
__global__ __launch_bounds__(256) void k_pos(BState st, uint32_t* keynew, const BScene* scenes,
                                             const BlockEntry* blocks, MoverOut mv,
                                             uint32_t* force, uint32_t* wdisp) {
    __shared__ uint32_t smax;
    const uint32_t L = xcd_remap(blockIdx.x, gridDim.x);
    const BlockEntry be = blocks[L];
    const BScene S = scenes[be.scene];
    if (force && threadIdx.x == 0) force[be.scene] = 1;
    if (wdisp) {  // uniform
        if (threadIdx.x == 0) smax = 0;
        __syncthreads();
    }
    const uint32_t j = be.j0 + threadIdx.x;
    float d = 0.0f;
    if (j < be.j1) {
        double4 p = st.pr[j];
        double x = p.x, y = p.y, z = p.z;
        if (ADV) {
            float4 vm = st.vm[j], aa = st.aa[j];
            x += S.dt * (double)vm.x + S.half2dt * (double)aa.x;
            y += S.dt * (double)vm.y + S.half2dt * (double)aa.y;
            z += S.dt * (double)vm.z + S.half2dt * (double)aa.z;
            st.pr[j] = make_double4(x, y, z, p.w);
            st.pf[j] = pf_of(S, x, y, z, (float)p.w);
            const float ddx = (float)(x - p.x), ddy = (float)(y - p.y), ddz = (float)(z - p.z);
            d = fsqrt(ddx * ddx + ddy * ddy + ddz * ddz) * 1.0001f;
        }
        int cx, cy, cz;
        cell_of(S, x, y, z, cx, cy, cz);
        const uint32_t kn = key_of(S, cx, cy, cz);
        keynew[j] = kn;
        if (mv.on) record_mover(mv, j, kn, st.zi[j].y, be.scene);
    }
    if (wdisp) {
        const uint32_t m = wave_max_hi16(__float_as_uint(d));  // every lane of the wave
        if (__lane_id() == 0) atomicMax(&smax, m);
        __syncthreads();
        if (threadIdx.x == 0) wdisp[L] = smax;
    }
}

// ------------------------------------------------------------------ sort
struct SortArgs {
    uint32_t n;                 // particles (storage positions)
    uint32_t ncell_cap;         // size of crange
    const uint32_t* key_old;    // sorted keys of the previous sort, storage order
    const uint2* zi;            // (alpha.z bits, particle id), storage order
    const uint32_t* patched;    // 1: k_patch already placed the movers (the kernels below exit)
    uint2* crange;
    uint64_t* moverbits;        // one bit per storage position (set by record_mover)
    uint32_t* blockcnt;         // movers per ROWS_T block; scanned in place (+1 total)
                                // when there are more than MERGE_LDS movers
    uint32_t* msorted;          // number of movers of this sort (written by k_patch)
    uint64_t* mkey;             // mover (key << 32 | id), append order
    uint32_t* ms;               // mover storage index, append order
    uint32_t* mplo;             // mover insertion point in the previous order, append order;
                                // second half: sorted mover storage indices (<= MERGE_LDS
                                // movers) or radix scratch
    uint64_t* mkey2;            // sorted movers (and radix ping-pong)
    uint32_t* ms2;              // sorted movers' append index
    uint32_t* perm;             // out: sorted position -> storage index
    uint32_t* invperm;          // out: storage index -> sorted position
    uint32_t* skey;             // the keys of the current positions, storage order (written by
                                // k_step / k_pos), sorted in place: stayers' keys are read from
                                // key_old, movers' from their records, never from here
    uint32_t nblk;              // ROWS_T blocks over n
    uint32_t key_bits, id_bits;
};

__device__ inline uint32_t wave_scan_incl(uint32_t v) {
    uint32_t lane = __lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Exclusive scan of the per-block mover counts in place (total at [nblk]),
// for k_merge's bit counts when more than MERGE_LDS movers; one block of NT.
template <int NT>
__device__ void scan_blockcnt(const SortArgs& A) {
    constexpr int NW = NT / 64;
    __shared__ uint32_t wsum[NW];
    __shared__ uint32_t carry_s;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < A.nblk; base += NT) {
        const uint32_t e = base + t;
        const uint32_t v = e < A.nblk ? A.blockcnt[e] : 0;
        const uint32_t inc = wave_scan_incl(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            const uint32_t x = lane < (uint32_t)NW ? wsum[lane] : 0;
            const uint32_t xi = wave_scan_incl(x);
            if (lane < (uint32_t)NW) wsum[lane] = xi - x;
        }
        __syncthreads();
        const uint32_t carry = carry_s;
        if (e < A.nblk) A.blockcnt[e] = carry + wsum[w] + inc - v;
        __syncthreads();
        if (t == NT - 1) carry_s = carry + wsum[w] + inc;
        __syncthreads();
    }
    if (t == 0) A.blockcnt[A.nblk] = carry_s;
}

// One workgroup of NT threads (block 0 of k_patch, beyond RANK_MAX movers:
// large admissions): the M movers sorted by (key, id) with an LSD radix sort
// in global memory, and the per-block mover counts scanned for k_merge.
template <int NT>
__device__ void sort_movers_block(const SortArgs& A, const uint32_t M) {
    constexpr int NW = NT / 64;
    __shared__ uint32_t lk[256];        // per-digit chunk totals
    __shared__ uint32_t wcnt[NW * 256]; // per wave and digit
    __shared__ uint32_t hist[256];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    scan_blockcnt<NT>(A);
    if (M == 0) return;
    // LSD radix over the id bits then the key bits; 8-bit digits; stable by
    // processing NT-item chunks in order.  Payload: the append index,
    // ping-ponged between the second half of the mplo buffer and ms2 in step
    // with the keys (mkey, mkey2).
    uint32_t* scratch = A.mplo + A.n;
    for (uint32_t i = t; i < M; i += NT) scratch[i] = i;
    __syncthreads();
    uint64_t* src_k = A.mkey;
    uint32_t* src_v = scratch;
    uint64_t* dst_k = A.mkey2;
    uint32_t* dst_v = A.ms2;
    int shifts[16];
    int npass = 0;
    for (uint32_t b = 0; b < A.id_bits; b += 8) shifts[npass++] = (int)b;
    for (uint32_t b = 0; b < A.key_bits; b += 8) shifts[npass++] = 32 + (int)b;
    for (int p = 0; p < npass; ++p) {
        int sh = shifts[p];
        if (t < 256) hist[t] = 0;
        __syncthreads();
        for (uint32_t i = t; i < M; i += NT) atomicAdd(&hist[(uint32_t)(src_k[i] >> sh) & 255u], 1u);
        __syncthreads();
        if (t == 0) {
            uint32_t run = 0;
            for (int d = 0; d < 256; ++d) {
                uint32_t c = hist[d];
                hist[d] = run;
                run += c;
            }
        }
        __syncthreads();
        const uint64_t lt = lanemask_lt64();
        for (uint32_t base = 0; base < M; base += NT) {
            for (uint32_t i = t; i < NW * 256; i += NT) wcnt[i] = 0;
            __syncthreads();
            uint32_t e = base + t;
            bool valid = e < M;
            uint64_t kk = valid ? src_k[e] : 0;
            uint32_t vv = valid ? src_v[e] : 0;
            uint32_t d = (uint32_t)(kk >> sh) & 255u;
            uint64_t active = __ballot(valid);
            uint64_t peers = active;
            for (int bb = 0; bb < 8; ++bb) {
                int bit = (d >> bb) & 1;
                uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            uint32_t rank = __popcll(peers & lt);
            if (valid && (peers & lt) == 0) wcnt[w * 256 + d] = __popcll(peers);
            __syncthreads();
            if (t < 256) {  // exclusive over waves per digit; chunk total of digit t
                uint32_t run = 0;
                for (int k = 0; k < NW; ++k) {
                    uint32_t c = wcnt[k * 256 + t];
                    wcnt[k * 256 + t] = run;
                    run += c;
                }
                lk[t] = run;
            }
            __syncthreads();
            if (valid) {
                uint32_t o = hist[d] + wcnt[w * 256 + d] + rank;
                dst_k[o] = kk;
                dst_v[o] = vv;
            }
            __syncthreads();
            if (t < 256) hist[t] += (uint32_t)lk[t];
            __syncthreads();
        }
        uint64_t* tk = src_k;
        src_k = dst_k;
        dst_k = tk;
        uint32_t* tv = src_v;
        src_v = dst_v;
        dst_v = tv;
        __syncthreads();
    }
    if (src_k != A.mkey2) {
        for (uint32_t i = t; i < M; i += NT) {
            A.mkey2[i] = src_k[i];
            A.ms2[i] = src_v[i];
        }
    }
}

// The M movers sorted by counting ranks (keys are distinct): groups of 64
// movers, one group per block at a time (grid-stride); lane l of the group's
// 64 holds mover l, and the block's T / 64 waves split the M keys between them
// (staged through LDS RANK_CHUNK at a time), so each thread counts M / (T / 64)
// keys; the partial counts are summed in LDS.  Mover r lands at its rank:
// mkey2 / ms2 (append index).  Every thread of the block calls it.
__device__ void rank_movers(const SortArgs& A, const uint32_t M) {
    __shared__ uint64_t cX[RANK_CHUNK];
    __shared__ uint32_t part[PATCH_T];
    const uint32_t t = threadIdx.x, T = blockDim.x, lane = t & 63, w = t >> 6, W = T >> 6;
    for (uint32_t g0 = blockIdx.x * 64; g0 < M; g0 += gridDim.x * 64) {  // block-uniform
        const uint32_t r = g0 + lane;
        const uint64_t myX = r < M ? A.mkey[r] : ~0ull;
        uint32_t rk = 0;
        for (uint32_t c0 = 0; c0 < M; c0 += RANK_CHUNK) {
            const uint32_t cn = min((uint32_t)RANK_CHUNK, M - c0);
            for (uint32_t i = t; i < cn; i += T) cX[i] = A.mkey[c0 + i];
            __syncthreads();
#pragma unroll 4
            for (uint32_t i = w; i < cn; i += W) rk += cX[i] < myX ? 1u : 0u;
            __syncthreads();
        }
        part[t] = rk;
        __syncthreads();
        if (w == 0 && r < M) {
            uint32_t sum = 0;
            for (uint32_t q = 0; q < W; ++q) sum += part[q * 64 + lane];
            A.mkey2[sum] = myX;
            A.ms2[sum] = r;
        }
        __syncthreads();
    }
}

// Movers at storage positions < P.
__device__ inline uint32_t mover_prefix(const SortArgs& A, uint32_t P) {
    if (P >= A.n) return A.blockcnt[A.nblk];
    uint32_t b = P / ROWS_T;
    uint32_t c = A.blockcnt[b];
    uint32_t w0 = b * (ROWS_T / 64), w = P >> 6;
    for (uint32_t k = w0; k < w; ++k) c += __popcll(A.moverbits[k]);
    uint32_t r = P & 63;
    if (r) c += __popcll(A.moverbits[w] & ((1ull << r) - 1));
    return c;
}

// Merge: stayers keep their relative order and shift by the movers placed
// before them; movers are placed by their rank among the previous order.
// Stayers count the sorted movers below their (key, id) by a binary search in
// LDS (global memory when more than MERGE_LDS movers).  A mover starts from
// its insertion point (recorded with it) and steps over the entries of its new cell
// with a smaller id.  Every placed element takes its position into its cell's
// range (atomic minimum of the first, maximum of one past the last; k_patch
// cleared the ranges of the previous order's cells).
__device__ inline uint32_t movers_below(const uint64_t* mk, uint32_t M, uint64_t X) {
    uint32_t lo = 0, hi = M;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (mk[mid] < X)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Grid: SWEEP_GRID blocks stepping over the ROWS_T tiles of stayers, then 64
// blocks for the movers.
__global__ __launch_bounds__(ROWS_T) void k_merge(SortArgs A) {
    __shared__ uint64_t smk[MERGE_LDS];
    if (*A.patched) return;
    const uint32_t M = *A.msorted;
    const bool few = M <= (uint32_t)MERGE_LDS;
    if (blockIdx.x < (uint32_t)SWEEP_GRID) {
        if (few)
            for (uint32_t i = threadIdx.x; i < M; i += ROWS_T) smk[i] = A.mkey2[i];
        __syncthreads();
        for (uint32_t tile = blockIdx.x; tile < A.nblk; tile += SWEEP_GRID) {
            const uint32_t s = tile * ROWS_T + threadIdx.x;
            if (s >= A.n) continue;
            // a mover (its bit was set by record_mover) is placed by the mover blocks
            if (M && ((A.moverbits[s >> 6] >> (s & 63)) & 1ull)) continue;
            const uint32_t kn = A.key_old[s];  // a stayer keeps its key
            uint32_t out = s;
            if (M) {
                const uint64_t X = ((uint64_t)kn << 32) | A.zi[s].y;
                const uint32_t lo = movers_below(few ? smk : A.mkey2, M, X);
                out = s - mover_prefix(A, s) + lo;
            }
            if (out < A.n) {
                A.perm[out] = s;
                A.invperm[s] = out;
                A.skey[out] = kn;
                if (kn < A.ncell_cap) {
                    atomicMin(&A.crange[kn].x, out);
                    atomicMax(&A.crange[kn].y, out + 1);
                }
            }
        }
        return;
    }
    const uint32_t nmb = gridDim.x - SWEEP_GRID;
    for (uint32_t r = (blockIdx.x - SWEEP_GRID) * ROWS_T + threadIdx.x; r < M; r += nmb * ROWS_T) {
        const uint64_t X = A.mkey2[r];
        const uint32_t p = A.ms2[r];
        if (p >= A.n) continue;  // defensive: payloads are append indices < M
        const uint32_t s = A.ms[p];
        const uint32_t kn = (uint32_t)(X >> 32);
        // entries of the previous order with (key, id) < X (recorded with the mover)
        const uint32_t lo = A.mplo[p];
        uint32_t out = r + (lo - mover_prefix(A, lo));
        if (out < A.n && s < A.n) {
            A.perm[out] = s;
            A.invperm[s] = out;
            A.skey[out] = kn;
            if (kn < A.ncell_cap) {
                atomicMin(&A.crange[kn].x, out);
                atomicMax(&A.crange[kn].y, out + 1);
            }
        }
    }
}

// Cell ranges of the sorted keys (reorderDataAndFindCellStart_kernel, :166-212)
// after a full sort or a layout change.  Grid-stride over the positions.
__global__ __launch_bounds__(ROWS_T) void k_ranges(const uint32_t* skey, uint2* crange, uint32_t n,
                                                  uint32_t ncell_cap) {
    for (uint32_t j = blockIdx.x * ROWS_T + threadIdx.x; j < n; j += gridDim.x * ROWS_T) {
        const uint32_t h = skey[j];
        if (h >= ncell_cap) continue;  // defensive: keys are < the cell capacity
        if (j == 0 || skey[j - 1] != h) {
            crange[h].x = j;
            if (j > 0 && skey[j - 1] < ncell_cap) crange[skey[j - 1]].y = j;
        }
        if (j == n - 1) crange[h].y = n;
    }
}

// ------------------------------------------------------- window patch
// The re-sort when at most FEW particles changed cell (the usual case between
// two steps).  With storage in the previous sorted order, a mover at storage
// src whose new (key, id) has insertion point lo in that order (the number of
// previous entries below it) shifts exactly the entries between src and lo,
// so the new order differs from the old one only inside the windows
// [src, lo - 1] (moving up) or [lo, src] (moving down), merged where they
// overlap or touch; every position outside them keeps its element, key and
// cell boundaries.  Each block recomputes the movers' order and windows (a
// wave-level sort of <= 64 records) and the grid strides over the window
// positions: position j's new element is found by inverting the merge (the
// mover placed at j, or the stayer of rank j - (movers placed before j) in the
// old order), and the cell boundaries at window positions are rewritten from
// the new keys.  Positions of the previous sort's windows outside the new
// ones get the identity permutation back.  More movers: they are ranked for
// k_merge, which does the rest (patched = 0).
constexpr int WIN_WORDS = 2 + 2 * FEW;   // per parity: count, then (first, last) pairs

struct PatchArgs {
    uint32_t n, ncell_cap, par;
    uint32_t patch_on;         // 0: every sort takes the full merge
    const uint32_t* key_old;   // sorted keys of the previous sort (storage order now)
    uint32_t* skey;            // this sort's keys, storage order (k_step / k_pos), patched in place
    const uint2* zi;           // particle ids, storage order
    uint2* crange;
    uint32_t* perm;
    uint32_t* invperm;
    uint32_t* counters;        // [0], [1] movers by key-buffer parity, [2] movers of the
                               // sort, [3] patched, [4], [5] scenes rebuilding every list
                               // by parity, [6], [7] the dirty tag of each parity's sort
    const uint64_t* mkey;      // mover records, append order (record_mover)
    const uint32_t* ms;
    const uint32_t* mplo;
    const uint4* mold;
    uint64_t* moverbits;
    uint32_t* blockcnt;
    uint32_t* win;             // windows by parity (this sort: par; the previous: 1 - par)
    // candidate lists (lists != 0)
    uint32_t lists;
    uint16_t* dirty;           // per cell key: tag of the sort that last marked it
    const BScene* scenes;
    uint32_t nscenes;
    float* disp;               // per scene: displacement bound since its lists were built
    const uint32_t* wdisp;     // per block of the last k_step: its largest displacement
    const uint2* scene_blk;    // per scene: first and end block of the launch table
    uint32_t* force;           // per scene: a block start asks for a rebuild
    uint32_t* rebuild;         // per scene: out, rebuild every list in the next step
    uint64_t* mb_clear;        // the other parity's mover marks and block counts, cleared
    uint32_t* bc_clear;        // here for the next sort
    uint32_t mb_words, bc_words;
};

// Cell `which` (0..26) of the 27 around cell key k of scene S (the walk's
// wrap-around) gets the sort's tag (plain stores of one value: marks of
// different movers never interfere).
__device__ inline void mark_neighbourhood(const BScene& S, uint32_t k, uint32_t which,
                                          uint16_t* dirty, uint16_t tag) {
    const uint32_t local = k - S.cell_base;
    const int x = (int)(local % S.gx), y = (int)((local / S.gx) % S.gy),
              z = (int)(local / (S.gx * S.gy));
    const int dz = (int)(which / 9) - 1, dy = (int)((which / 3) % 3) - 1, dx = (int)(which % 3) - 1;
    const uint32_t kk = key_of(S, x + dx, y + dy, z + dz);
    dirty[kk] = tag;
}

// Inclusive scans of a wave's 64 lanes by DPP moves: shifts by 1, 2, 4 and 8
// lanes inside each row of 16, then lane 15 of a row added into the next row
// (rows 1 and 3) and lane 31 into the upper half (rows 2 and 3).  Lanes with
// no source take 0, the identity of both operations.
__device__ inline uint32_t wave_incl_sum(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)x;
}
__device__ inline uint32_t wave_incl_max(uint32_t v) {
    int x = (int)v;
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
    x = (int)max((uint32_t)x, (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
    return (uint32_t)x;
}

// Block-wide scans over PATCH_T threads (one value per thread): inclusive
// maximum and exclusive sums; every thread of the block calls them.
__device__ inline uint32_t block_incl_max(uint32_t v, uint32_t* wtmp) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_incl_max(v);
    if (lane == 63) wtmp[w] = v;
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PATCH_T / 64; ++q) {
        const uint32_t u = wtmp[q];
        if (q < w) v = max(v, u);
    }
    __syncthreads();
    return v;
}
__device__ inline uint32_t block_excl_sum(uint32_t v, uint32_t* wtmp, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t x = wave_incl_sum(v);
    if (lane == 63) wtmp[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t q = 0; q < PATCH_T / 64; ++q) {
        const uint32_t u = wtmp[q];
        before += q < w ? u : 0u;
        all += u;
    }
    __syncthreads();
    total = all;
    return before + x - v;
}

__device__ inline uint2 block_excl_sum2(uint2 v, uint32_t* wtmp, uint2& total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint2 x = make_uint2(wave_incl_sum(v.x), wave_incl_sum(v.y));
    if (lane == 63) {
        wtmp[2 * w] = x.x;
        wtmp[2 * w + 1] = x.y;
    }
    __syncthreads();
    uint2 before = make_uint2(0u, 0u), all = make_uint2(0u, 0u);
#pragma unroll
    for (uint32_t q = 0; q < PATCH_T / 64; ++q) {
        const uint32_t a = wtmp[2 * q], b = wtmp[2 * q + 1];
        if (q < w) {
            before.x += a;
            before.y += b;
        }
        all.x += a;
        all.y += b;
    }
    __syncthreads();
    total = all;
    return make_uint2(before.x + x.x - v.x, before.y + x.y - v.y);
}

// Entries of the ascending v[0..m) (stride st) below x / at most x.
__device__ inline uint32_t lds_count_lt(const uint32_t* v, uint32_t m, uint32_t st, uint32_t x) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v[mid * st] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
__device__ inline uint32_t lds_count_le(const uint32_t* v, uint32_t m, uint32_t st, uint32_t x) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v[mid * st] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
__device__ inline uint32_t lds_count_lt64(const uint64_t* v, uint32_t m, uint64_t x) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Ranks among the keys of threads 0..M-1 (M <= PATCH_T) in three key sets at
// once: the movers' (key, id), their storage indices and the windows' (start,
// append index).  Each set holds distinct keys, so no tie is tested.  The keys
// form chunks of 64 (by thread); every wave of the block takes a share of the
// comparisons: first each chunk's 64 x 64 (its keys' ranks inside the chunk;
// the chunk is then written in that order to cb*), then, per pair of chunks,
// a binary search of each key of one chunk in the other (the three searches
// in step).  Partial counts are summed with LDS atomics in R*.
// kb*, cb*, R*: PATCH_T entries of LDS each.  Every thread of the block calls
// it; the buffers may be reused after the next barrier.
__device__ inline uint3 block_rank(uint64_t k1, uint32_t k2, uint64_t k3, uint32_t M,
                                   uint64_t* kb1, uint64_t* cb1, uint32_t* kb2, uint32_t* cb2,
                                   uint64_t* kb3, uint64_t* cb3, uint32_t* R1, uint32_t* R2,
                                   uint32_t* R3) {
    constexpr uint32_t NWV = PATCH_T / 64;
    const uint32_t t = threadIdx.x, lane = t & 63;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(t >> 6));
    const bool has = t < M;
    if (has) {
        kb1[t] = k1;
        kb2[t] = k2;
        kb3[t] = k3;
    }
    R1[t] = 0;
    R2[t] = 0;
    R3[t] = 0;
    __syncthreads();
    const uint32_t nch = (M + 63) >> 6;
    // ---- inside each chunk: S waves per chunk, each a slice of its 64 keys
    const uint32_t S = nch ? max(1u, NWV / nch) : 1u;
    {
        const uint32_t c = w / S, part = w - c * S;
        if (c < nch) {
            const uint32_t b0 = c * 64, cn = min(64u, M - b0);
            const uint32_t e = b0 + lane;  // the key this lane ranks
            const bool mine = lane < cn;
            const uint64_t x1 = mine ? kb1[e] : 0ull;
            const uint32_t x2 = mine ? kb2[e] : 0u;
            const uint64_t x3 = mine ? kb3[e] : 0ull;
            const uint32_t r0 = part * 64 / S, r1e = min((part + 1) * 64 / S, cn);
            uint32_t a1 = 0, a2 = 0, a3 = 0;
            // the same entries for every lane (broadcasts), four in flight
            uint32_t r = r0;
            for (; r + 4 <= r1e; r += 4) {
                uint64_t o1[4], o3[4];
                uint32_t o2[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    o1[u] = kb1[b0 + r + u];
                    o2[u] = kb2[b0 + r + u];
                    o3[u] = kb3[b0 + r + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    a1 += o1[u] < x1 ? 1u : 0u;
                    a2 += o2[u] < x2 ? 1u : 0u;
                    a3 += o3[u] < x3 ? 1u : 0u;
                }
            }
            for (; r < r1e; ++r) {
                a1 += kb1[b0 + r] < x1 ? 1u : 0u;
                a2 += kb2[b0 + r] < x2 ? 1u : 0u;
                a3 += kb3[b0 + r] < x3 ? 1u : 0u;
            }
            if (mine && r0 < r1e) {
                atomicAdd(&R1[e], a1);
                atomicAdd(&R2[e], a2);
                atomicAdd(&R3[e], a3);
            }
        }
    }
    __syncthreads();
    const uint32_t c0 = w * 64;
    if (has) {
        cb1[c0 + R1[t]] = k1;
        cb2[c0 + R2[t]] = k2;
        cb3[c0 + R3[t]] = k3;
    }
    __syncthreads();
    // ---- across chunks: pair q = (chunk of the keys, chunk searched)
    for (uint32_t q = w; q < nch * nch; q += NWV) {
        const uint32_t ce = q / nch, cs = q - ce * nch;
        if (ce == cs) continue;
        const uint32_t e = ce * 64 + lane;
        if (e >= M) continue;
        const uint64_t x1 = kb1[e], x3 = kb3[e];
        const uint32_t x2 = kb2[e];
        const uint32_t b0 = cs * 64, m = min(64u, M - b0);
        uint32_t lo1 = 0, hi1 = m, lo2 = 0, hi2 = m, lo3 = 0, hi3 = m;
        while (lo1 < hi1 || lo2 < hi2 || lo3 < hi3) {  // the three searches in step
            if (lo1 < hi1) {
                const uint32_t mid = (lo1 + hi1) >> 1;
                if (cb1[b0 + mid] < x1)
                    lo1 = mid + 1;
                else
                    hi1 = mid;
            }
            if (lo2 < hi2) {
                const uint32_t mid = (lo2 + hi2) >> 1;
                if (cb2[b0 + mid] < x2)
                    lo2 = mid + 1;
                else
                    hi2 = mid;
            }
            if (lo3 < hi3) {
                const uint32_t mid = (lo3 + hi3) >> 1;
                if (cb3[b0 + mid] < x3)
                    lo3 = mid + 1;
                else
                    hi3 = mid;
            }
        }
        if (lo1) atomicAdd(&R1[e], lo1);
        if (lo2) atomicAdd(&R2[e], lo2);
        if (lo3) atomicAdd(&R3[e], lo3);
    }
    __syncthreads();
    return has ? make_uint3(R1[t], R2[t], R3[t]) : make_uint3(0u, 0u, 0u);
}

// Each block works out the movers' order and the merged windows in LDS, thread
// t holding mover t of the append order: ranks by (key, id) and by storage
// index (block_rank), each mover's new position (its rank plus the previous
// entries below its insertion point, less the movers among them) and its
// window between old and new position, the windows ranked by (first, last)
// and merged by a block-wide running maximum of their last positions.  More
// than FEW movers, or new windows holding more than PATCH_SLOTS_MAX positions
// (movers across far-apart cells), take the full merge instead.  Roles then:
// block 0 the counters, block 1 the emptied cells, the movers' marks and the
// window record, blocks 2.. the sweep from the bottom (new window positions,
// one past each too for its last boundary, then the previous windows'
// positions outside the new ones) and the dirty marks from the top; then
// every block a share of the scenes' displacement bounds.  A position's
// element comes from binary searches: the mover placed there, or the stayer
// whose rank among the stayers is j - (movers placed before j).  On the full
// merge the blocks sort the movers for k_merge instead.
constexpr uint32_t PATCH_WSLOTS = 63;  // sweep positions per wave: lane 0 works out the one before
constexpr uint32_t PATCH_SLOTS_MAX = (PATCH_GRID - 2u) * (PATCH_T / 64) * PATCH_WSLOTS;

__global__ __launch_bounds__(PATCH_T) void k_patch(PatchArgs A, SortArgs SA) {
    __shared__ uint64_t sX[FEW];                  // by rank: (key, id)
    __shared__ uint32_t sSrc[FEW], sOut[FEW];     // by rank: storage index, new position
    __shared__ uint32_t sScene[FEW];              // by rank: scene
    __shared__ uint32_t sAsc[FEW];                // storage indices ascending
    __shared__ uint64_t kb1[FEW], cb1[FEW];       // block_rank buffers
    __shared__ uint32_t kb2[FEW], cb2[FEW];
    __shared__ uint64_t kb3[FEW], cb3[FEW];
    __shared__ uint32_t R1[FEW], R2[FEW], R3[FEW];  // block_rank counts
    __shared__ uint64_t wk[FEW];                  // windows (first << 32 | last) in order
    __shared__ uint32_t sWin[2 * FEW];            // merged windows: (first, last) pairs, ordered
    __shared__ uint32_t sPrev[2 * FEW];           // the previous sort's windows
    __shared__ uint32_t sWpre[FEW + 1], sPpre[FEW + 1];  // slots before each window
    __shared__ uint32_t wtmp[2 * (PATCH_T / 64)];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t n = A.n;
    uint32_t* const win_now = A.win + A.par * WIN_WORDS;
    const uint32_t* const win_prev = A.win + (1 - A.par) * WIN_WORDS;
    const uint32_t M = min(A.counters[A.par], n);
    const uint32_t npv = min(win_prev[0], (uint32_t)FEW);
    // this sort's dirty marks carry a new tag (the previous sort's plus one,
    // 1 .. 65535): the next k_step takes the cells holding it as marked, so
    // older marks need no clearing
    const uint32_t tag = A.counters[6 + (1 - A.par)] % 65535u + 1u;
    bool few = A.patch_on && M <= (uint32_t)FEW;
    // the previous sort's mover marks and block counts, cleared for the next
    for (uint32_t q = blockIdx.x * PATCH_T + t; q < A.mb_words; q += gridDim.x * PATCH_T)
        A.mb_clear[q] = 0;
    for (uint32_t q = blockIdx.x * PATCH_T + t; q < A.bc_words; q += gridDim.x * PATCH_T)
        A.bc_clear[q] = 0;
    if (t < npv) {
        sPrev[2 * t] = win_prev[2 + 2 * t];
        sPrev[2 * t + 1] = win_prev[3 + 2 * t];
    }
    // this thread's mover (append index t), one memory round trip
    const bool rec = few && t < M;
    const uint64_t mk = rec ? A.mkey[t] : 0ull;
    const uint32_t src = rec ? A.ms[t] : 0u;
    const uint32_t ins = rec ? A.mplo[t] : 0u;
    const uint4 mo = rec ? A.mold[t] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t nw = 0;
    if (few && M > 0) {  // (no movers: no new windows)
        // the mover's window: the positions between its old storage index and
        // its insertion point in the previous order; windows go in start order
        // (equal starts by append index: the merge below does not depend on
        // their order)
        const uint32_t wa = src < ins ? src : ins, wb = src < ins ? ins - 1 : src;
        const uint64_t wkey = ((uint64_t)wa << 32) | wb;
        const uint64_t wrank = ((uint64_t)wa << 32) | t;
        const uint3 rk = block_rank(mk, src, wrank, M, kb1, cb1, kb2, cb2, kb3, cb3, R1, R2, R3);
        if (rec) {
            sX[rk.x] = mk;
            sSrc[rk.x] = src;
            sScene[rk.x] = mo.y;
            sAsc[rk.y] = src;
            wk[rk.z] = wkey;
        }
        __syncthreads();
        if (rec) {
            const uint32_t below = lds_count_lt(sAsc, M, 1, ins);
            sOut[rk.x] = rk.x + ins - below;
        }
        // merged where they overlap or touch: a running maximum of the last
        // positions decides where a merged window starts and ends
        const bool valid = t < M;
        const uint64_t wv = valid ? wk[t] : 0ull;
        const uint32_t a = valid ? (uint32_t)(wv >> 32) : 0xFFFFFFFFu;
        const uint32_t b = valid ? (uint32_t)wv : 0u;
        const uint32_t run = block_incl_max(b, wtmp);
        uint32_t* const wA = kb2;  // free again
        uint32_t* const wB = cb2;
        if (valid) wA[t] = run;
        __syncthreads();
        const bool start = valid && (t == 0 || a > wA[t - 1] + 1);
        if (valid) wB[t] = start ? 1u : 0u;
        uint32_t nstart = 0;
        const uint32_t k = block_excl_sum(start ? 1u : 0u, wtmp, nstart);
        const bool next_start = t + 1 < M && wB[t + 1] != 0;
        if (start) sWin[2 * k] = a;
        if (valid && (t + 1 == M || next_start)) sWin[2 * (k + (start ? 1u : 0u) - 1) + 1] = run;
        nw = nstart;
    }
    __syncthreads();
    // ---- slot prefixes of the new windows (one past each too) and of the
    // previous ones: the sweep finds a slot's window by search
    uint32_t total_new = 0, total_prev = 0;
    if (nw | npv) {
        uint2 tot;
        const uint2 e = block_excl_sum2(
            make_uint2(t < nw ? sWin[2 * t + 1] - sWin[2 * t] + 2 : 0u,
                       t < npv ? sPrev[2 * t + 1] - sPrev[2 * t] + 1 : 0u), wtmp, tot);
        if (t < nw) sWpre[t] = e.x;
        if (t < npv) sPpre[t] = e.y;
        if (t == 0) {
            sWpre[nw] = tot.x;
            sPpre[npv] = tot.y;
        }
        total_new = tot.x;
        total_prev = tot.y;
        __syncthreads();
    }
    if (few && total_new > PATCH_SLOTS_MAX) few = false;  // the same decision in every block
    if (blockIdx.x == 0 && t == 0) {
        A.counters[1 - A.par] = 0;  // the next step's movers go there
        A.counters[2] = M;
        A.counters[3] = few ? 1u : 0u;
        A.counters[4 + (1 - A.par)] = 0;
        A.counters[6 + A.par] = tag;
        if (!few) {  // the full merge rewrites every position
            win_now[0] = 1;
            win_now[2] = 0;
            win_now[3] = n - 1;
        }
    }
    if (!few) {
        // the full merge follows: the movers are sorted for it — by counting
        // ranks spread over the blocks (64 movers per block at a time), or
        // beyond RANK_MAX by one block (large admissions)
        if (M <= (uint32_t)RANK_MAX) {
            // the last block scans the per-block mover counts for k_merge
            if (blockIdx.x == gridDim.x - 1) scan_blockcnt<PATCH_T>(SA);
            rank_movers(SA, M);
        } else if (blockIdx.x == 0) {
            sort_movers_block<PATCH_T>(SA, M);
        }
        // and the previous order's cell ranges are cleared (k_merge writes
        // every cell's range of the new order)
        for (uint32_t j = blockIdx.x * PATCH_T + t; j < n; j += gridDim.x * PATCH_T) {
            const uint32_t k = A.key_old[j];
            if (k < A.ncell_cap && (j == 0 || A.key_old[j - 1] != k))
                A.crange[k] = make_uint2(EMPTY, 0);
        }
    } else if (blockIdx.x == 1) {
        // ---- a cell every particle left (all its positions carry a mover
        // mark) and none entered becomes empty; the windows of this sort are
        // recorded
        if (rec) {
            const uint32_t ko = mo.z, size = mo.x, first = mo.w;
            if (ko < A.ncell_cap && size > 0) {
                bool all = true;
                for (uint32_t p = first; all && p < first + size;) {
                    const uint32_t b0 = p & 63, nb = min(64u - b0, first + size - p);
                    const uint64_t mask = (nb == 64 ? ~0ull : ((1ull << nb) - 1ull)) << b0;
                    all = (A.moverbits[p >> 6] & mask) == mask;
                    p += nb;
                }
                if (all && lds_count_lt64(sX, M, (uint64_t)(ko + 1) << 32) ==
                               lds_count_lt64(sX, M, (uint64_t)ko << 32))
                    A.crange[ko] = make_uint2(EMPTY, 0);
            }
        }
        if (t < nw) {
            win_now[2 + 2 * t] = sWin[2 * t];
            win_now[3 + 2 * t] = sWin[2 * t + 1];
        }
        if (t == 0) win_now[0] = nw;
    } else if (blockIdx.x >= 2) {
        // ---- dirty cells (blocks from the top): the 27 cells around each
        // mover's new cell get this sort's tag (the movers' records are at
        // hand on the patch path only; otherwise every list is rebuilt)
        if (A.lists)
            for (uint32_t q = (gridDim.x - 1 - blockIdx.x) * PATCH_T + t; q < 27 * M;
                 q += (gridDim.x - 2) * PATCH_T) {
                const uint32_t e = q / 27;
                mark_neighbourhood(A.scenes[sScene[e]], (uint32_t)(sX[e] >> 32), q % 27, A.dirty,
                                   (uint16_t)tag);
            }
        // ---- the sweep (blocks from the bottom), PATCH_WSLOTS positions per
        // wave: lane 0 works out the element before the wave's first position
        // and every lane passes its key to the next, which needs it for the
        // cell boundary.  The keys around a position are loaded before its
        // element is worked out (a stayer in a window comes from at most one
        // position away per overlapping mover, usually one).
        auto elem = [&](uint32_t j, const uint4 q, uint32_t& key) -> uint32_t {
            const uint32_t c1 = lds_count_lt(sOut, M, 1, j);  // movers placed before j
            if (c1 < M && sOut[c1] == j) {
                key = (uint32_t)(sX[c1] >> 32);
                return sSrc[c1];
            }
            // the stayer of rank q0 = j - c1: q0 plus the movers before it, i.e.
            // those whose storage index exceeds the stayers before it by at most
            // q0 (sAsc[i] - i, the stayers before mover i, does not decrease)
            const uint32_t q0 = j - c1;
            uint32_t lo = 0, hi = M;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sAsc[mid] - mid <= q0)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            const uint32_t s = min(q0 + lo, n - 1);  // defensive: the windows hold their own stayers
            if (s + 2 >= j && s <= j + 1) {
                const uint32_t d = s + 2 - j;
                key = d == 0 ? q.x : d == 1 ? q.y : d == 2 ? q.z : q.w;
            } else {
                key = A.key_old[s];
            }
            return s;
        };
        // the previous order's keys at j + d (outside every window a position
        // keeps its key)
        auto kat = [&](uint32_t j, int d) -> uint32_t {
            const int64_t y = (int64_t)j + d;
            return y >= 0 && y < (int64_t)n ? A.key_old[(uint32_t)y] : EMPTY;
        };
        const uint32_t total = total_new + total_prev;
        constexpr uint32_t BS = (PATCH_T / 64) * PATCH_WSLOTS;
        for (uint32_t g0 = (blockIdx.x - 2) * BS; g0 < total; g0 += (gridDim.x - 2) * BS) {
            const uint32_t gw = g0 + w * PATCH_WSLOTS;  // the wave's first slot
            const uint32_t g = gw + lane - 1;           // lane 0: the slot before (not written)
            const bool real = lane != 0 && g < total;
            const bool seen = (lane != 0 || gw != 0) && g < total_new;
            uint32_t j = 0, a = 0, kc = EMPTY;
            uint4 q = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
            bool inr = false;
            if (seen) {
                const uint32_t c = lds_count_le(sWpre, nw, 1, g) - 1;  // the window holding slot g
                a = sWin[2 * c];
                const uint32_t b = sWin[2 * c + 1];
                j = a + (g - sWpre[c]);
                inr = j <= n;
                if (inr) {
                    q = make_uint4(kat(j, -2), kat(j, -1), kat(j, 0), kat(j, 1));
                    if (j <= b) {
                        const uint32_t e = elem(j, q, kc);
                        if (real) {
                            A.perm[j] = e;
                            A.invperm[e] = j;
                            A.skey[j] = kc;
                        }
                    } else if (j < n) {
                        kc = q.z;  // one past the window: unchanged
                    }
                }
            }
            // the key before j: the lane before worked out slot g - 1, which is
            // position j - 1 of the same window when j - 1 >= a
            const uint32_t kn = (uint32_t)__shfl_up((int)kc, 1, 64);
            if (!real) continue;
            if (g < total_new) {
                if (!inr) continue;
                const uint32_t kp = j == 0 ? EMPTY : (j - 1 >= a ? kn : q.y);
                if (j == 0 || j == n || kp != kc) {
                    if (j < n && kc < A.ncell_cap) A.crange[kc].x = j;
                    if (j > 0 && kp < A.ncell_cap) A.crange[kp].y = j;
                }
            } else {
                const uint32_t gp = g - total_new;
                const uint32_t c = lds_count_le(sPpre, npv, 1, gp) - 1;
                const uint32_t jj = sPrev[2 * c] + (gp - sPpre[c]);
                // inside a new window? (merged windows are disjoint and ordered)
                const uint32_t cn = lds_count_le(sWin, nw, 2, jj);
                const bool in_new = cn > 0 && jj <= sWin[2 * (cn - 1) + 1];
                if (jj < n && !in_new) {
                    A.perm[jj] = jj;
                    A.invperm[jj] = jj;
                }
            }
        }
    }
    // ---- per scene (blocks in turn): the displacement bound since its lists
    // were built, the step's largest block value added; every list is rebuilt
    // when it reaches half the skin (both particles of a pair may have moved),
    // after a contact kernel without lists or a position update that did not
    // record its displacement (force), or on the full merge
    if (A.lists)
        for (uint32_t sc = blockIdx.x; sc < A.nscenes; sc += gridDim.x) {
            const uint2 br = A.scene_blk[sc];
            uint32_t m = 0;
            for (uint32_t q0 = br.x + t; q0 < br.y; q0 += 4 * PATCH_T) {
                uint32_t v[4];  // four loads in flight
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t q = q0 + u * PATCH_T;
                    v[u] = q < br.y ? A.wdisp[q] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) m = max(m, v[u]);
            }
            for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
            if (lane == 0) wtmp[w] = m;
            __syncthreads();
            if (t == 0) {
                for (uint32_t q = 0; q < PATCH_T / 64; ++q) m = max(m, wtmp[q]);
                const float D = A.disp[sc] + __uint_as_float(m);
                const bool rb = A.force[sc] != 0 || !few || !(2.0f * D < A.scenes[sc].skin);
                A.disp[sc] = rb ? 0.0f : D;
                A.rebuild[sc] = rb ? 1u : 0u;
                A.force[sc] = 0;
                if (rb) atomicAdd(&A.counters[4 + A.par], 1u);
            }
            __syncthreads();
        }
}

// Reset of the sort and list state after a layout change or a regather:
// both mover counters zero, both window sets the whole range (the next patch
// restores the identity everywhere), every scene's lists rebuilt (the dirty
// tags and the mover marks are cleared by the host).
// One block.
__global__ __launch_bounds__(256) void k_sort_reset(PatchArgs A) {
    const uint32_t t = threadIdx.x;
    if (t < 2) {
        A.counters[t] = 0;
        A.counters[4 + t] = 0;
        A.win[t * WIN_WORDS] = A.n ? 1u : 0u;
        A.win[t * WIN_WORDS + 2] = 0;
        A.win[t * WIN_WORDS + 3] = A.n ? A.n - 1 : 0u;
    }
    if (A.lists)
        for (uint32_t sc = t; sc < A.nscenes; sc += 256) {
            A.force[sc] = 1;
            A.disp[sc] = 0.0f;
            A.rebuild[sc] = 1;
        }
}

// invperm[perm[j]] = j (after a full sort).
__global__ void k_invert(const uint32_t* perm, uint32_t* invperm, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const uint32_t s = perm[j];
        if (s < n) invperm[s] = j;
    }
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_fill_u2(uint2* p, uint2 v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

}  // namespace fused
}  // namespace xdem
