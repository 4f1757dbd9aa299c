// xdem_setb.hip — the older entry points of the reference
// (include/cu_dem_dec.cuh:95-113; include/xdem_setb.h): double4 layouts
// with acc.w = mass and alpha.w = inertia, constants in the parameter block.
//
// FP64 kernels, thread per item, 256-thread blocks.  Each kernel takes the
// number of ids it processes, min(T(n), bound) (the reference's computeGridSize
// thread count capped by the parameter-block bound, cu_dem_def.cu:13-23), so
// the reference's launch-bound behaviour is kept exactly.  Vector division is
// a multiply by the reciprocal as the reference's operator/
// (cu_dem_impl.cuh:68-71).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/xdem_setb.h"
#include "xdem_common.hpp"

static_assert(sizeof(xdem_b_device_parameters) == 144, "device_parameters (cu_dem_dec.cuh:41-60)");
static_assert(sizeof(xdem_b_plane_info) == 160, "device_plane_info (cu_dem_dec.cuh:62-71)");
static_assert(sizeof(xdem_b_polygon) == 144, "cu_polygon (cu_dem_dec.cuh:73-81)");
static_assert(sizeof(xdem_b_mass_info) == 48, "cu_mass_info (cu_dem_dec.cuh:83-87)");

namespace xdem {
namespace setb {

using P = xdem_b_device_parameters;

xdem_b_device_parameters g_params;
bool g_set = false;

constexpr uint32_t T = 256;

// computeGridSize(n, 256) thread count, capped by the kernel's bound.
uint32_t ids_below(uint32_t n, uint64_t bound) {
    if (n == 0) return 0;
    uint64_t t = std::min<uint32_t>(T, n);
    uint64_t threads = ((n + t - 1) / t) * t;
    return (uint32_t)std::min(threads, bound);
}

__device__ inline d3 ld4(const double* p) { return mk3(p[0], p[1], p[2]); }
__device__ inline d3 ld3(const double* p) { return mk3(p[0], p[1], p[2]); }
__device__ inline void st3(double* p, d3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ inline d3 divr(d3 a, double b) {  // cu_dem_impl.cuh:68-71
    double inv = 1 / b;
    return mk3(a.x * inv, a.y * inv, a.z * inv);
}
__device__ inline double len(d3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ inline int sgn(double a) { return a < 0 ? -1 : 1; }  // :10-13

// calcGridPos / calcGridHash — cu_dem_impl.cuh:73-91
__device__ inline void cell_of(const P& p, d3 q, int& cx, int& cy, int& cz) {
    cx = (int)floor((q.x - p.world_origin.x) / p.cell_size);
    cy = (int)floor((q.y - p.world_origin.y) / p.cell_size);
    cz = (int)floor((q.z - p.world_origin.z) / p.cell_size);
}
__device__ inline uint32_t hash_of(const P& p, int cx, int cy, int cz) {
    uint32_t x = (uint32_t)cx & (p.grid_size.x - 1);
    uint32_t y = (uint32_t)cy & (p.grid_size.y - 1);
    uint32_t z = (uint32_t)cz & (p.grid_size.z - 1);
    return __umul24(__umul24(z, p.grid_size.y), p.grid_size.x) + __umul24(y, p.grid_size.x) + x;
}

__global__ __launch_bounds__(T) void kb_vv_position(P p, double* pos, const double* vel,
                                                    const double* acc, uint32_t n) {
    uint32_t i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    d3 q = ld4(pos + 4 * (size_t)i);
    q += p.dt * ld4(vel + 4 * (size_t)i) + p.half2dt * ld4(acc + 4 * (size_t)i);
    st3(pos + 4 * (size_t)i, q);
}

__global__ __launch_bounds__(T) void kb_vv_velocity(P p, double* vel, double* acc, double* omega,
                                                    double* alpha, const double* force,
                                                    const double* moment, uint32_t n) {
    uint32_t i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    const size_t q = 4 * (size_t)i;
    d3 v = ld4(vel + q), a = ld4(acc + q), av = ld4(omega + q), aa = ld4(alpha + q);
    double mass = acc[q + 3], inertia = alpha[q + 3];
    v += 0.5 * p.dt * a;
    av += 0.5 * p.dt * aa;
    a = (1 / mass) * ld3(force + 3 * (size_t)i);
    aa = (1 / inertia) * ld3(moment + 3 * (size_t)i);
    v += 0.5 * p.dt * a;
    av += 0.5 * p.dt * aa;
    st3(vel + q, v);
    vel[q + 3] = 0;
    st3(omega + q, av);
    omega[q + 3] = 0;
    st3(acc + q, a);
    acc[q + 3] = mass;
    st3(alpha + q, aa);
    alpha[q + 3] = inertia;
}

__global__ __launch_bounds__(T) void kb_mergedata(double* data, const double* pos, uint32_t np,
                                                  const double* vert, uint32_t n) {
    uint32_t i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    const size_t q = 4 * (size_t)i;
    if (i < np) {
        data[q] = pos[q];
        data[q + 1] = pos[q + 1];
        data[q + 2] = pos[q + 2];
        data[q + 3] = pos[q + 3];
    } else {
        const double* v = vert + 3 * (size_t)(i - np);
        data[q] = v[0];
        data[q + 1] = v[1];
        data[q + 2] = v[2];
        data[q + 3] = -1.0;
    }
}

__global__ __launch_bounds__(T) void kb_hash(P p, uint32_t* hash, uint32_t* index, const double* pos,
                                             uint32_t n) {
    uint32_t i = blockIdx.x * T + threadIdx.x;
    if (i >= n) return;
    int cx, cy, cz;
    cell_of(p, ld4(pos + 4 * (size_t)i), cx, cy, cz);
    hash[i] = hash_of(p, cx, cy, cz);
    index[i] = i;
}

// reorderDataAndFindCellStart_kernel — cu_dem_impl.cuh:156-196 (the shared
// hash[id - 1] read directly from memory)
__global__ __launch_bounds__(T) void kb_reorder(const uint32_t* hash, const uint32_t* index,
                                                uint32_t* cstart, uint32_t* cend, uint32_t* sorted,
                                                uint32_t n, uint64_t tnp) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= n) return;
    uint32_t h = hash[id];
    if (id == 0 || h != hash[id - 1]) {
        cstart[h] = id;
        if (id > 0) cend[hash[id - 1]] = id;
    }
    if ((uint64_t)id == tnp - 1) cend[h] = id + 1;
    sorted[id] = index[id];
}

// calForce — cu_dem_impl.cuh:198-252
__device__ bool cal_force(const P& c, double ir, double jr, d3 ipos, d3 jpos, d3 ivel, d3 jvel,
                          d3 iomega, d3 jomega, d3& force, d3& moment) {
    d3 rp = jpos - ipos;
    double dist = len(rp);
    double cd = (ir + jr) - dist;
    if (cd <= 0) {
        if (c.cohesive == 0) return false;
        double tu = c.cohesive / c.mu;
        d3 unit = divr(rp, dist);
        d3 rv = jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit));
        double Fn = (-c.kn * pow(-cd, 1.5) + c.vn * dot(rv, unit));
        if (Fn < 0 && Fn > -tu) force += tu * unit;
        d3 e = rv - dot(rv, unit) * unit;
        double mag_e = len(e);
        if (mag_e) {
            d3 s_hat = divr(e, mag_e);
            double ds = mag_e * c.dt;
            double Fs = fmin(c.ks * ds + c.vs * (dot(rv, s_hat)), c.mu * Fn);
            if (Fs < Fn * c.mu + c.cohesive) moment += cross(ir * unit, (Fn * c.mu + c.cohesive) * s_hat);
        }
        return false;
    }
    d3 unit = divr(rp, dist);
    d3 rv = jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit));
    d3 single_force = (-c.kn * pow(cd, 1.5) + c.vn * dot(rv, unit)) * unit;
    d3 single_moment = mk3(0, 0, 0);
    d3 e = rv - dot(rv, unit) * unit;
    double mag_e = len(e);
    if (mag_e) {
        d3 s_hat = divr(e, mag_e);
        double ds = mag_e * c.dt;
        d3 shear = fmin(c.ks * ds + c.vs * (dot(rv, s_hat)), c.mu * len(single_force)) * s_hat;
        single_moment = cross(ir * unit, shear);
    }
    force += single_force;
    moment += single_moment;
    return true;
}

// calculate_p2p_kernel — cu_dem_impl.cuh:254-327
__global__ __launch_bounds__(T) void kb_p2p(P p, const double* pos, const double* vel,
                                            const double* acc, const double* omega, double* force,
                                            double* moment, const uint32_t* sorted,
                                            const uint32_t* cstart, const uint32_t* cend,
                                            uint32_t n) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= n) return;
    d3 ipos = ld4(pos + 4 * (size_t)id), ivel = ld4(vel + 4 * (size_t)id),
       iomega = ld4(omega + 4 * (size_t)id);
    double ir = pos[4 * (size_t)id + 3];
    int gx, gy, gz;
    cell_of(p, ipos, gx, gy, gz);
    d3 F = acc[4 * (size_t)id + 3] * of(p.gravity), M = mk3(0, 0, 0);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t h = hash_of(p, gx + x, gy + y, gz + z);
                uint32_t s = cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = sorted[j];
                    if (id == k || k >= p.np) continue;
                    cal_force(p, ir, pos[4 * (size_t)k + 3], ipos, ld4(pos + 4 * (size_t)k), ivel,
                              ld4(vel + 4 * (size_t)k), iomega, ld4(omega + 4 * (size_t)k), F, M);
                }
            }
    double* f = force + 3 * (size_t)id;
    double* m = moment + 3 * (size_t)id;
    f[0] += F.x;
    f[1] += F.y;
    f[2] += F.z;
    m[0] += M.x;
    m[1] += M.y;
    m[2] += M.z;
}

struct WallK {
    double kn, vn, ks, vs, mu;
};

// cube_hertzian_contact_force_kernel — cu_dem_impl.cuh:329-378
__global__ __launch_bounds__(T) void kb_cube(P p, const xdem_b_plane_info* planes, WallK w,
                                             const double* pos, const double* vel,
                                             const double* omega, double* force, double* moment,
                                             uint32_t n) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= n) return;
    double r = pos[4 * (size_t)id + 3];
    d3 ipos = ld4(pos + 4 * (size_t)id), ivel = ld4(vel + 4 * (size_t)id),
       iomega = ld4(omega + 4 * (size_t)id);
    d3 mf = mk3(0, 0, 0), mm = mk3(0, 0, 0);
    for (int i = 0; i < 6; i++) {
        const xdem_b_plane_info pl = planes[i];
        d3 dp = ipos - of(pl.xw);
        d3 wp = mk3(dot(dp, of(pl.u1)), dot(dp, of(pl.u2)), dot(dp, of(pl.uw)));
        if (fabs(wp.z) < r && (wp.x > 0 && wp.x < pl.l1) && (wp.y > 0 && wp.y < pl.l2)) {
            d3 unit = (double)(-sgn(dot(ipos - of(pl.xw), of(pl.uw)))) * divr(of(pl.uw), len(of(pl.uw)));
            double cd = r - fabs(dot(ipos - of(pl.xw), unit));
            d3 dv = -(ivel + cross(iomega, r * unit));
            d3 sf = (-w.kn * pow(cd, 1.5) + w.vn * dot(dv, unit)) * unit;
            d3 e = dv - dot(dv, unit) * unit;
            double mag_e = len(e);
            if (mag_e) {
                d3 s_hat = divr(e, mag_e);
                double ds = mag_e * p.dt;
                d3 shear = fmin(w.ks * ds + w.vs * dot(dv, s_hat), w.mu * len(sf)) * s_hat;
                mm += cross(r * unit, shear);
            }
            mf += sf;
        }
    }
    double* f = force + 3 * (size_t)id;
    double* m = moment + 3 * (size_t)id;
    f[0] += mf.x;
    f[1] += mf.y;
    f[2] += mf.z;
    m[0] += mm.x;
    m[1] += mm.y;
    m[2] += mm.z;
}

// ClosestPtPointTriangle with region code — cu_dem_impl.cuh:380-441
__device__ d3 closest_pt(d3 q, d3 a, d3 b, d3 c, int* wc) {
    d3 ab = b - a, ac = c - a, ap = q - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        *wc = 0;
        return a;
    }
    d3 bp = q - b;
    double d3_ = dot(ab, bp), d4 = dot(ac, bp);
    if (d3_ >= 0.0 && d4 <= d3_) {
        *wc = 0;
        return b;
    }
    double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) {
        *wc = 1;
        double v = d1 / (d1 - d3_);
        return a + v * ab;
    }
    d3 cp = q - c;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        *wc = 0;
        return c;
    }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        *wc = 1;
        double w = d2 / (d2 - d6);
        return a + w * ac;
    }
    double va = d3_ * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3_) >= 0.0 && (d5 - d6) >= 0.0) {
        *wc = 1;
        double w = (d4 - d3_) / ((d4 - d3_) + (d5 - d6));
        return b + w * (c - b);
    }
    *wc = 2;
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom;
    double w = vc * denom;
    return a + v * ab + w * ac;
}

struct ShapeArgs {
    P p;
    const xdem_b_mass_info* mi;
    const xdem_b_polygon* polys;
    const uint32_t *id_set, *poly_start, *poly_end;
    WallK w;
    bool* is_line;
    const double *pos, *vel, *omega;
    double *force, *moment;
    const uint32_t *sorted, *cstart, *cend;
    uint32_t n;
};

// Face pass (shpae_hertzian_contact_force_kernel :676-762 with
// calculateTriangleParticleContactForce :454-526) when FACE, else the
// line/edge pass (shape_line_or_edge_contact_force_kernel :591-674 with
// :528-589): the first qualifying triangle contact in traversal order.  The
// face pass returns at its first contact, so the line intersection test of
// :494-500 (which needs an earlier contact of the same call) never runs.
template <bool FACE>
__global__ __launch_bounds__(T) void kb_shape(ShapeArgs A) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= A.n) return;
    const P& c = A.p;
    if (FACE)
        A.is_line[id] = false;
    else if (A.is_line[id])
        return;
    d3 ipos = ld4(A.pos + 4 * (size_t)id), ivel = ld4(A.vel + 4 * (size_t)id),
       iomega = ld4(A.omega + 4 * (size_t)id);
    double ir = A.pos[4 * (size_t)id + 3];
    d3 bvel = of(A.mi->vel);
    int gx, gy, gz;
    cell_of(c, ipos, gx, gy, gz);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t h = hash_of(c, gx + x, gy + y, gz + z);
                uint32_t s = A.cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = A.cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = A.sorted[j];
                    if (id == k || k < c.np) continue;
                    uint32_t vid = k - c.np;
                    for (uint32_t i = A.poly_start[vid]; i < A.poly_end[vid]; ++i) {
                        const xdem_b_polygon& po = A.polys[A.id_set[i]];
                        int wc = -1;
                        d3 cp = closest_pt(ipos, of(po.P), of(po.Q), of(po.R), &wc);
                        d3 sp = ipos - cp;
                        double dist = len(sp);
                        double cd = ir - dist;
                        if (cd <= 0) continue;
                        if (FACE ? wc != 2 : (wc != 0 && wc != 1)) continue;
                        d3 unit = FACE ? divr(-of(po.N), len(of(po.N))) : divr(-sp, dist);
                        d3 dv = bvel - (ivel + cross(iomega, ir * unit));
                        d3 sf = (-A.w.kn * pow(cd, 1.5) + A.w.vn * dot(dv, unit)) * unit;
                        d3 mo = mk3(0, 0, 0);
                        d3 et = dv - dot(dv, unit) * unit;
                        double mag_e = len(et);
                        if (mag_e) {
                            d3 s_hat = divr(et, mag_e);
                            double ds = mag_e * c.dt;
                            d3 shear = fmin(A.w.ks * ds + A.w.vs * dot(dv, s_hat), A.w.mu * len(sf)) * s_hat;
                            mo = cross(ir * unit, shear);
                        }
                        double* f = A.force + 3 * (size_t)id;
                        double* m = A.moment + 3 * (size_t)id;
                        f[0] += sf.x;
                        f[1] += sf.y;
                        f[2] += sf.z;
                        m[0] += mo.x;
                        m[1] += mo.y;
                        m[2] += mo.z;
                        A.is_line[id] = true;
                        return;
                    }
                }
            }
}

// reduce6<double3, 256> — cu_dem_impl.cuh:820-847, launched with
// computeGridSize(n, 256) blocks; lanes at or beyond nthreads contribute zero.
__global__ __launch_bounds__(T) void kb_reduce256(const d3* in, d3* out, uint32_t n,
                                                  uint32_t nthreads) {
    __shared__ d3 sdata[T];
    uint32_t tid = threadIdx.x;
    uint64_t i = (uint64_t)blockIdx.x * (T * 2) + tid;
    uint64_t grid = (uint64_t)T * 2 * gridDim.x;
    d3 s = mk3(0.0, 0.0, 0.0);
    if (tid < nthreads) {
        while (i < n) {
            s += in[i];
            if (i + T < n) s += in[i + T];
            i += grid;
        }
    }
    sdata[tid] = s;
    __syncthreads();
    for (uint32_t h = T / 2; h >= 1; h >>= 1) {
        if (tid < h) sdata[tid] = s = s + sdata[tid + h];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(T) void kb_fill_f64(double* p, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * T + threadIdx.x;
    size_t stp = (size_t)gridDim.x * T;
    for (; i < n; i += stp) p[i] = v;
}

// update_polygons_position — cu_dem_impl.cuh:764-780
__global__ __launch_bounds__(T) void kb_poly_position(const double* mpos, const double* A,
                                                      const double* lvx, uint32_t nv,
                                                      double* pos) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= nv) return;
    d3 lv = ld3(lvx + 3 * (size_t)id);
    double tx = A[0] * lv.x + A[1] * lv.y + A[2] * lv.z;
    double ty = A[3] * lv.x + A[4] * lv.y + A[5] * lv.z;
    double tz = A[6] * lv.x + A[7] * lv.y + A[8] * lv.z;
    pos[3 * (size_t)id] = mpos[0] + tx;
    pos[3 * (size_t)id + 1] = mpos[1] + ty;
    pos[3 * (size_t)id + 2] = mpos[2] + tz;
}

// update_polygons_information — cu_dem_impl.cuh:798-818
__global__ __launch_bounds__(T) void kb_poly_info(xdem_b_polygon* polys, const uint32_t* ind,
                                                  const double* pos, uint32_t ni) {
    uint32_t id = blockIdx.x * T + threadIdx.x;
    if (id >= ni) return;
    const uint32_t* ix = ind + 3 * (size_t)id;
    d3 Pp = ld3(pos + 3 * (size_t)ix[0]), Q = ld3(pos + 3 * (size_t)ix[1]),
       R = ld3(pos + 3 * (size_t)ix[2]);
    d3 V = Q - Pp, W = R - Pp, N = cross(V, W);
    xdem_b_polygon po;
    po.P = xdem_double3{Pp.x, Pp.y, Pp.z};
    po.Q = xdem_double3{Q.x, Q.y, Q.z};
    po.R = xdem_double3{R.x, R.y, R.z};
    po.V = xdem_double3{V.x, V.y, V.z};
    po.W = xdem_double3{W.x, W.y, W.z};
    po.N = xdem_double3{N.x, N.y, N.z};
    polys[id] = po;
}

int32_t need() {
    if (!g_set) {
        set_error("xdem_b_set_symbolic_parameter has not been called");
        return XDEM_ERR_STATE;
    }
    return XDEM_OK;
}

inline dim3 grid_for(uint32_t n) { return dim3(div_up(n, T)); }

}  // namespace setb
}  // namespace xdem

using namespace xdem;
using namespace xdem::setb;

extern "C" {

int32_t xdem_b_set_symbolic_parameter(const xdem_b_device_parameters* p) {
    XDEM_REQUIRE(p, "null parameters");
    XDEM_REQUIRE(p->cell_size > 0, "cell_size must be > 0");
    for (uint32_t g : {p->grid_size.x, p->grid_size.y, p->grid_size.z})
        XDEM_REQUIRE(g && !(g & (g - 1)), "grid_size must be powers of two");
    g_params = *p;
    g_set = true;
    return XDEM_OK;
}

int32_t xdem_b_get_symbolic_parameter(xdem_b_device_parameters* p) {
    XDEM_REQUIRE(p, "null parameters");
    XDEM_TRY(need());
    *p = g_params;
    return XDEM_OK;
}

int32_t xdem_b_vv_update_position(double* pos4, const double* vel4, const double* acc4,
                                  uint32_t np) {
    XDEM_TRY(need());
    uint32_t n = ids_below(np, g_params.np);
    if (!n) return XDEM_OK;
    XDEM_REQUIRE(pos4 && vel4 && acc4, "null array");
    hipLaunchKernelGGL(kb_vv_position, grid_for(n), dim3(T), 0, stream(), g_params, pos4, vel4,
                       acc4, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_vv_update_velocity(double* vel4, double* acc4, double* omega4, double* alpha4,
                                  const double* force3, const double* moment3, uint32_t np) {
    XDEM_TRY(need());
    uint32_t n = ids_below(np, g_params.np);
    if (!n) return XDEM_OK;
    XDEM_REQUIRE(vel4 && acc4 && omega4 && alpha4 && force3 && moment3, "null array");
    hipLaunchKernelGGL(kb_vv_velocity, grid_for(n), dim3(T), 0, stream(), g_params, vel4, acc4,
                       omega4, alpha4, force3, moment3, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_mergedata(double* data4, const double* pos4, uint32_t np, const double* vertice3,
                         uint32_t snp) {
    uint64_t tot = (uint64_t)np + snp;
    XDEM_REQUIRE(tot < 0xFFFFFFFFull, "too many items");
    if (!tot) return XDEM_OK;
    XDEM_REQUIRE(data4 && (np == 0 || pos4) && (snp == 0 || vertice3), "null array");
    hipLaunchKernelGGL(kb_mergedata, grid_for((uint32_t)tot), dim3(T), 0, stream(), data4, pos4, np,
                       vertice3, (uint32_t)tot);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_calculate_hash_and_index(uint32_t* hash, uint32_t* index, const double* pos4,
                                        uint32_t np) {
    XDEM_TRY(need());
    if (!np) return XDEM_OK;
    XDEM_REQUIRE(hash && index && pos4, "null array");
    uint32_t n = ids_below(np, (uint64_t)g_params.np + g_params.nsp);
    if (n) {
        hipLaunchKernelGGL(kb_hash, grid_for(n), dim3(T), 0, stream(), g_params, hash, index, pos4,
                           n);
        XDEM_LAUNCHED();
    }
    // thrust::sort_by_key over the np entries (cu_dem_def.cu:66-68): stable
    uint64_t cells = (uint64_t)g_params.grid_size.x * g_params.grid_size.y * g_params.grid_size.z;
    uint32_t bits = cells > (1ull << 24) ? 32u : bits_for((uint32_t)cells);
    return sort_pairs(hash, index, np, bits, stream());
}

int32_t xdem_b_reorder_data_and_find_cell_start(const uint32_t* hash, const uint32_t* index,
                                                uint32_t* cstart, uint32_t* cend,
                                                uint32_t* sorted_index, uint32_t np,
                                                uint32_t nc) {
    XDEM_TRY(need());
    XDEM_REQUIRE(nc == 0 || cstart, "null cstart");
    if (nc) XDEM_TRY(fill_u32(cstart, 0xFFFFFFFFu, nc, stream()));
    const uint64_t tnp = (uint64_t)g_params.np + g_params.nsp;
    uint32_t n = ids_below(np, tnp);
    if (!n) return XDEM_OK;
    XDEM_REQUIRE(hash && index && cend && sorted_index, "null array");
    hipLaunchKernelGGL(kb_reorder, grid_for(n), dim3(T), 0, stream(), hash, index, cstart, cend,
                       sorted_index, n, tnp);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_calculate_p2p(const double* pos4, const double* vel4, const double* acc4,
                             const double* omega4, const double* alpha4, double* force3,
                             double* moment3, const uint32_t* sorted_index,
                             const uint32_t* cstart, const uint32_t* cend, uint32_t np,
                             uint32_t cRun) {
    (void)alpha4;
    (void)cRun;
    XDEM_TRY(need());
    uint32_t n = ids_below(np, g_params.np);
    if (!n) return XDEM_OK;
    XDEM_REQUIRE(pos4 && vel4 && acc4 && omega4 && force3 && moment3 && sorted_index && cstart &&
                     cend,
                 "null array");
    hipLaunchKernelGGL(kb_p2p, grid_for(n), dim3(T), 0, stream(), g_params, pos4, vel4, acc4,
                       omega4, force3, moment3, sorted_index, cstart, cend, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_cube_hertzian_contact_force(const xdem_b_plane_info* planes6, double kn,
                                           double vn, double ks, double vs, double mu,
                                           bool* is_line_contact, const double* pos4,
                                           const double* vel4, const double* omega4,
                                           double* force3, double* moment3, uint32_t np) {
    (void)is_line_contact;
    XDEM_TRY(need());
    uint32_t n = ids_below(np, g_params.np);
    if (!n) return XDEM_OK;
    XDEM_REQUIRE(planes6 && pos4 && vel4 && omega4 && force3 && moment3, "null array");
    hipLaunchKernelGGL(kb_cube, grid_for(n), dim3(T), 0, stream(), g_params, planes6,
                       WallK{kn, vn, ks, vs, mu}, pos4, vel4, omega4, force3, moment3, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_shape_hertzian_contact_force(const xdem_b_mass_info* minfo,
                                            const xdem_b_polygon* polygons,
                                            const uint32_t* id_set, const uint32_t* poly_start,
                                            const uint32_t* poly_end, double kn, double vn,
                                            double ks, double vs, double mu,
                                            bool* is_line_contact, const double* pos4,
                                            const double* vel4, const double* omega4,
                                            double* force3, double* moment3,
                                            double* body_force3, uint32_t np,
                                            const uint32_t* sorted_id,
                                            const uint32_t* cell_start,
                                            const uint32_t* cell_end, double out_body_force[3]) {
    (void)body_force3;
    XDEM_TRY(need());
    XDEM_REQUIRE(out_body_force, "null result");
    out_body_force[0] = out_body_force[1] = out_body_force[2] = 0.0;
    if (!np) return XDEM_OK;
    XDEM_REQUIRE(force3 && moment3, "null array");
    hipStream_t s = stream();
    const size_t nd = 3 * (size_t)np;
    // cudaMemset(force/moment, 0) of cu_dem_def.cu:166-167, as a fill kernel
    const uint32_t fb = (uint32_t)std::min<size_t>(1024, (nd + T - 1) / T);
    hipLaunchKernelGGL(kb_fill_f64, dim3(fb), dim3(T), 0, s, force3, 0.0, nd);
    XDEM_LAUNCHED();
    hipLaunchKernelGGL(kb_fill_f64, dim3(fb), dim3(T), 0, s, moment3, 0.0, nd);
    XDEM_LAUNCHED();
    uint32_t n = ids_below(np, g_params.np);
    if (n) {
        XDEM_REQUIRE(minfo && polygons && id_set && poly_start && poly_end && is_line_contact &&
                         pos4 && vel4 && omega4 && sorted_id && cell_start && cell_end,
                     "null array");
        ShapeArgs A{g_params, minfo, polygons, id_set, poly_start, poly_end,
                    WallK{kn, vn, ks, vs, mu}, is_line_contact, pos4, vel4, omega4, force3,
                    moment3, sorted_id, cell_start, cell_end, n};
        hipLaunchKernelGGL(kb_shape<true>, grid_for(n), dim3(T), 0, s, A);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(kb_shape<false>, grid_for(n), dim3(T), 0, s, A);
        XDEM_LAUNCHED();
    }
    // reduce6<double3, 256> of the forces, block partials summed in order
    uint32_t nthreads = std::min<uint32_t>(T, np);
    uint32_t nblocks = div_up(np, nthreads);
    XDEM_TRY(workspace(3).reserve(sizeof(d3) * nblocks));
    d3* part = static_cast<d3*>(workspace(3).ptr);
    hipLaunchKernelGGL(kb_reduce256, dim3(nblocks), dim3(T), 0, s,
                       reinterpret_cast<const d3*>(force3), part, np, nthreads);
    XDEM_LAUNCHED();
    std::vector<d3> h(nblocks);
    XDEM_HIP(hipMemcpyAsync(h.data(), part, sizeof(d3) * nblocks, hipMemcpyDeviceToHost, s));
    XDEM_HIP(hipStreamSynchronize(s));
    double rx = 0, ry = 0, rz = 0;
    for (uint32_t i = 0; i < nblocks; ++i) {
        rx += h[i].x;
        ry += h[i].y;
        rz += h[i].z;
    }
    out_body_force[0] = rx;
    out_body_force[1] = ry;
    out_body_force[2] = rz;
    return XDEM_OK;
}

int32_t xdem_b_update_polygons_position(const double* mpos3, const double* A9,
                                        const double* local_vertice3, uint32_t nvertex,
                                        double* pos3) {
    if (!nvertex) return XDEM_OK;
    XDEM_REQUIRE(mpos3 && A9 && local_vertice3 && pos3, "null array");
    hipLaunchKernelGGL(kb_poly_position, grid_for(nvertex), dim3(T), 0, stream(), mpos3, A9,
                       local_vertice3, nvertex, pos3);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t xdem_b_update_polygons_information(xdem_b_polygon* polygons, const uint32_t* indice3,
                                           const double* pos3, uint32_t ni) {
    if (!ni) return XDEM_OK;
    XDEM_REQUIRE(polygons && indice3 && pos3, "null array");
    hipLaunchKernelGGL(kb_poly_info, grid_for(ni), dim3(T), 0, stream(), polygons, indice3, pos3,
                       ni);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

}  // extern "C"
