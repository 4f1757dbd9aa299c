// xdem_f64.hip — the plain FP64 path: one kernel per reference kernel, thread
// per particle, the reference AoS layout (pos double4, vectors packed double3)
// and its block sizes (512, 256 for the contact kernels).  This is the drop-in
// of the live Set A API (include/mphysics_cuda_dec.cuh:110-161) and the FP64
// comparison point of the fused FP32c path.
//
// Arithmetic restates include/mphysics_cuda_impl.cuh expression by expression;
// the library is compiled with -ffp-contract=off so nothing is fused into FMA
// (the oracle is compiled the same way).  Transcendentals (pow, log, sqrt of
// the force constants, atan/sin/cos of the cylinder rim) come from the ROCm
// device library and may differ from glibc by an ulp.
#include <algorithm>
#include <vector>

#include "xdem_common.hpp"
#include "xdem_f64.hpp"
#include "xdem_geom.hpp"

namespace xdem {

struct ForceConst {
    double kn, vn, ks, vs, mu, ms;
};

// getConstant — mphysics_cuda_impl.cuh:214-256
template <int TCM>
__device__ inline ForceConst get_constant(double ir, double jr, double im, double jm,
                                          const xdem_device_contact_property& cp) {
    ForceConst dfc = {0, 0, 0, 0, 0, 0};
    double Meq = jm ? (im * jm) / (im + jm) : im;
    double Req = jr ? (ir * jr) / (ir + jr) : ir;
    double Eeq = (cp.Ei * cp.Ej) / (cp.Ei * (1 - cp.prj * cp.prj) + cp.Ej * (1 - cp.pri * cp.pri));
    if (TCM == 0) {
        double Geq = (cp.Gi * cp.Gj) / (cp.Gi * (2 - cp.prj) + cp.Gj * (2 - cp.pri));
        double ln_e = log(cp.rest);
        double xi = ln_e / sqrt(ln_e * ln_e + M_PI * M_PI);
        dfc.kn = (4.0 / 3.0) * Eeq * sqrt(Req);
        dfc.vn = -2.0 * sqrt(5.0 / 6.0) * xi * sqrt(dfc.kn * Meq);
        dfc.ks = 8.0 * Geq * sqrt(Req);
        dfc.vs = -2.0 * sqrt(5.0 / 6.0) * xi * sqrt(dfc.ks * Meq);
    } else {
        double beta = (M_PI / log(cp.rest));
        dfc.kn = (4.0 / 3.0) * Eeq * sqrt(Req);
        dfc.vn = sqrt((4.0 * Meq * dfc.kn) / (1.0 + beta * beta));
        dfc.ks = dfc.kn * cp.sratio;
        dfc.vs = dfc.vn * cp.sratio;
    }
    dfc.mu = cp.fric;
    dfc.ms = cp.rfric;
    return dfc;
}

// cohesionForce — mphysics_cuda_impl.cuh:259-279 (particle-particle only; a
// wall/triangle contact has no cohesion term instead of the reference's 0/0).
__device__ inline double cohesion_force(double ri, double rj, double Ei, double Ej, double pri,
                                        double prj, double coh, double Fn) {
    double cf = 0.0;
    if (coh) {
        double req = (ri * rj / (ri + rj));
        double Eeq = ((1.0 - pri * pri) / Ei) + ((1.0 - prj * prj) / Ej);
        double rcp = (3.0 * req * (-Fn)) / (4.0 * (1.0 / Eeq));
        double rc = pow(rcp, 1.0 / 3.0);
        double Ac = M_PI * rc * rc;
        cf = coh * Ac;
    }
    return cf;
}

// DHSModel (:379-401) / HMCModel (:347-377).  Ft and M keep the caller's
// values when the tangential velocity is exactly zero, as in the reference.
template <int TCM>
__device__ inline void force_law(const ForceConst& c, double dt, double ir, double fca,
                                 double rcon, double cdist, d3 iomega, d3 dv, d3 unit, d3& Ft,
                                 d3& Fn, d3& M) {
    double fsn = -c.kn * pow(cdist, 1.5);
    double fdn = c.vn * dot(dv, unit);
    Fn = (fsn + fca + fdn) * unit;
    d3 e = dv - dot(dv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        if (TCM == 0) {
            d3 s_hat = -(e / mag_e);
            double ds = mag_e * dt;
            double fst = -c.ks * ds;
            double fdt = c.vs * dot(dv, s_hat);
            Ft = (fst + fdt) * s_hat;
            if (length(Ft) >= c.mu * length(Fn)) Ft = c.mu * fsn * s_hat;
            M = cross(ir * unit, Ft);
            if (length(iomega)) {
                d3 on = iomega / length(iomega);
                M += c.ms * fsn * rcon * on;
            }
        } else {
            d3 sh = e / mag_e;
            double ds = mag_e * dt;
            Ft = fmin(c.ks * ds + c.vs * (dot(dv, sh)), c.mu * length(Fn)) * sh;
            M = cross(ir * unit, Ft);
        }
    }
}

namespace {

__device__ inline d3 ld3(const double* p, uint32_t i) {
    const double* q = p + 3 * (size_t)i;
    return d3{q[0], q[1], q[2]};
}
__device__ inline void st3(double* p, uint32_t i, d3 v) {
    double* q = p + 3 * (size_t)i;
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
}
__device__ inline void add3(double* p, uint32_t i, d3 v) {
    double* q = p + 3 * (size_t)i;
    q[0] += v.x;
    q[1] += v.y;
    q[2] += v.z;
}
__device__ inline d4 ld4(const double* p, uint32_t i) {
    return reinterpret_cast<const d4*>(p)[i];
}

// K1 vv_update_position_kernel — mphysics_cuda_impl.cuh:85-96
__global__ __launch_bounds__(512) void k_vv_position(KParams k, double* pos, const double* vel,
                                                     const double* acc, uint32_t np) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    d3 p = k.dt * ld3(vel, id) + k.half2dt * ld3(acc, id);
    double* q = pos + 4 * (size_t)id;
    q[0] += p.x;
    q[1] += p.y;
    q[2] += p.z;
}

// K2 vv_update_velocity_kernel — mphysics_cuda_impl.cuh:98-134
__global__ __launch_bounds__(512) void k_vv_velocity(KParams k, double* vel, double* acc,
                                                     double* omega, double* alpha, double* force,
                                                     double* moment, const double* mass,
                                                     const double* iner, uint32_t np) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    double m = mass[id];
    d3 v = ld3(vel, id);
    d3 av = ld3(omega, id);
    d3 a = (1.0 / m) * ld3(force, id);
    d3 in = (1.0 / iner[id]) * ld3(moment, id);
    v += 0.5 * k.dt * (ld3(acc, id) + a);
    av += 0.5 * k.dt * (ld3(alpha, id) + in);
    st3(force, id, m * mk3(k.gxv, k.gyv, k.gzv));
    st3(moment, id, mk3(0.0, 0.0, 0.0));
    st3(vel, id, v);
    st3(omega, id, av);
    st3(acc, id, a);
    st3(alpha, id, in);
}

// K3 / K4 calculateHashAndIndex(_ForPolygonSphere)_kernel — :137-164
__global__ __launch_bounds__(512) void k_hash(KParams k, uint32_t* hash, uint32_t* index,
                                              const double* pos, uint32_t sid, uint32_t n) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    d4 p = ld4(pos, id);
    int cx, cy, cz;
    grid_pos(k, p.x, p.y, p.z, cx, cy, cz);
    hash[sid + id] = grid_hash(k, cx, cy, cz);
    index[sid + id] = sid + id;
}

// FP32 screening record of a particle: its position relative to the world
// origin and its radius, x NaN when a coordinate lies farther than pf_range
// from the origin (such a particle is never screened out: NaN distances fail
// every rejection).  Each coordinate is off by at most 2^-24 pf_range.
__device__ inline float4 pf_screen(const KParams& k, double x, double y, double z, double r,
                                   double pf_range) {
    const double rx = x - k.wox, ry = y - k.woy, rz = z - k.woz;
    const bool in = fabs(rx) <= pf_range && fabs(ry) <= pf_range && fabs(rz) <= pf_range;
    return make_float4(in ? (float)rx : __builtin_nanf(""), (float)ry, (float)rz, (float)r);
}

// The sorted screening copy for k_p2p: per sorted position the record above;
// polygon spheres (ids >= np, skipped by the particle pass) far away.
__global__ __launch_bounds__(256) void k_sorted_screen(KParams k, const double* pos,
                                                       const uint32_t* sorted, uint32_t np,
                                                       uint32_t nsorted, double pf_range,
                                                       float4* pfs) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsorted) return;
    const uint32_t id = sorted[j];
    float4 f = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    if (id < np) {
        const d4 p = ld4(pos, id);
        f = pf_screen(k, p.x, p.y, p.z, p.w, pf_range);
    }
    pfs[j] = f;
}

// K8 calculate_p2p_kernel<TCM> — mphysics_cuda_impl.cuh:403-485.  SORTED: thread t
// evaluates the particle at sorted position t (of nsorted entries, triangle
// spheres included and skipped) instead of particle t, so a wave's particles
// are neighbours in space and share their cells, and with pfs (the sorted
// screening copy, error bound pf_eps = 2^-20 pf_range on distances) screens
// each candidate before loading it; each particle's arithmetic and neighbour
// order are the same either way.
template <int TCM, bool SORTED>
__global__ __launch_bounds__(256) void k_p2p(KParams k, const double* pos, const double* vel,
                                             const double* omega, double* force, double* moment,
                                             const double* mass, const uint32_t* sorted_index,
                                             const uint32_t* cstart, const uint32_t* cend,
                                             const xdem_device_contact_property* cpp, uint32_t np,
                                             uint32_t nsorted, const float4* pfs, double pf_range,
                                             float pf_eps) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t id = t;
    if (SORTED) {
        if (t >= nsorted) return;
        id = sorted_index[t];
    }
    if (id >= np) return;
    const xdem_device_contact_property cp = *cpp;
    d4 ip4 = ld4(pos, id);
    d3 ipos = mk3(ip4.x, ip4.y, ip4.z);
    d3 ivel = ld3(vel, id), iomega = ld3(omega, id);
    int gx, gy, gz;
    grid_pos(k, ipos.x, ipos.y, ipos.z, gx, gy, gz);
    double ir = ip4.w, im = mass[id];
    d3 Ft = mk3(0, 0, 0), Fn = mk3(0, 0, 0), M = mk3(0, 0, 0);
    d3 sumF = mk3(0, 0, 0), sumM = mk3(0, 0, 0);
    // with the sorted screening copy: a candidate whose FP32 distance exceeds
    // ir + jr by more than twice its error bound cannot touch (the FP64 test
    // below would reject it), so neither its index nor its record is loaded
    const float4 ownf = pf_screen(k, ipos.x, ipos.y, ipos.z, ir, pf_range);
    const float sr0 = ownf.w + pf_eps;
    // candidate at sorted position j
    auto visit = [&](uint32_t j) {
        if (SORTED && pfs) {
            const float4 f = pfs[j];
            const float dx = f.x - ownf.x, dy = f.y - ownf.y, dz = f.z - ownf.z;
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const float sr = sr0 + f.w;
            if (d2 >= sr * sr * (1.0f + 0x1p-19f)) return;
        }
        uint32_t kk = sorted_index[j];
        if (id == kk || kk >= np) return;
        d4 jp = ld4(pos, kk);
        d3 rp = mk3(jp.x - ipos.x, jp.y - ipos.y, jp.z - ipos.z);
        double dist = length(rp);
        double jr = jp.w;
        double cdist = (ir + jr) - dist;
        if (cdist > 0) {
            d3 jvel = ld3(vel, kk), jomega = ld3(omega, kk);
            double jm = mass[kk];
            double rcon = ir - 0.5 * cdist;
            d3 unit = rp / dist;
            d3 rv = k.rolling ? jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit))
                              : (jvel - ivel);
            ForceConst c = get_constant<TCM>(ir, jr, im, jm, cp);
            double fsn = -c.kn * pow(cdist, 1.5);
            double fca = cohesion_force(ir, jr, cp.Ei, cp.Ej, cp.pri, cp.prj, cp.coh, fsn);
            force_law<TCM>(c, k.dt, ir, fca, rcon, cdist, iomega, rv, unit, Ft, Fn, M);
            sumF += Fn + Ft;
            sumM += M;
        }
    };
    if (SORTED && gx >= 1 && gx + 1 < (int)k.gx) {
        // Away from the x wrap the three cells of an x-row have consecutive
        // hashes, so their particles are one run of sorted positions: the
        // row's cells in x order, each in ascending sorted position — the
        // cell loop's order.  The six range words of a row are loaded together.
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++) {
                const uint32_t h = grid_hash(k, gx, gy + y, gz + z);
                const uint32_t s0 = cstart[h - 1], s1 = cstart[h], s2 = cstart[h + 1];
                const uint32_t e0 = cend[h - 1], e1 = cend[h], e2 = cend[h + 1];
                const uint32_t rb = s0 != 0xFFFFFFFFu ? s0 : (s1 != 0xFFFFFFFFu ? s1 : s2);
                if (rb == 0xFFFFFFFFu) continue;
                const uint32_t re = s2 != 0xFFFFFFFFu ? e2 : (s1 != 0xFFFFFFFFu ? e1 : e0);
                for (uint32_t j = rb; j < re; j++) visit(j);
            }
    } else {
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    uint32_t h = grid_hash(k, gx + x, gy + y, gz + z);
                    uint32_t s = cstart[h];
                    if (s == 0xFFFFFFFFu) continue;
                    uint32_t e = cend[h];
                    for (uint32_t j = s; j < e; j++) visit(j);
                }
    }
    add3(force, id, sumF);
    add3(moment, id, sumM);
}

// Contact pair list of K8's predicate: partners of each particle in traversal
// order (z, y, x cells; ascending sorted position), at most max_per each.
__global__ __launch_bounds__(256) void k_pairs(KParams k, const double* pos,
                                               const uint32_t* sorted_index, const uint32_t* cstart,
                                               const uint32_t* cend, uint32_t np, uint32_t max_per,
                                               uint32_t* counts, uint32_t* partners) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    d4 ip4 = ld4(pos, id);
    int gx, gy, gz;
    grid_pos(k, ip4.x, ip4.y, ip4.z, gx, gy, gz);
    uint32_t c = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t h = grid_hash(k, gx + x, gy + y, gz + z);
                uint32_t s = cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = cend[h];
                for (uint32_t j = s; j < e; j++) {
                    uint32_t kk = sorted_index[j];
                    if (id == kk || kk >= np) continue;
                    d4 jp = ld4(pos, kk);
                    d3 rp = mk3(jp.x - ip4.x, jp.y - ip4.y, jp.z - ip4.z);
                    double cdist = (ip4.w + jp.w) - length(rp);
                    if (cdist > 0) {
                        if (c < max_per) partners[(size_t)id * max_per + c] = kk;
                        ++c;
                    }
                }
            }
    counts[id] = c;
}

}  // namespace

namespace {

// K9 plane_contact_force_kernel<TCM> — mphysics_cuda_impl.cuh:572-621.  One
// launch applies `nplane` planes in order (the reference launches once per
// plane; each adds into force/moment, so the per-particle sums are the same).
template <int TCM>
__global__ __launch_bounds__(256) void k_plane(KParams k, const xdem_device_plane_info* planes,
                                               uint32_t nplane, const double* pos, const double* vel,
                                               const double* omega, double* force, double* moment,
                                               const xdem_device_contact_property* cpp,
                                               const double* mass, uint32_t np) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    double m = mass[id];
    d4 ip4 = ld4(pos, id);
    d3 ipos3 = mk3(ip4.x, ip4.y, ip4.z);
    double r = ip4.w;
    d3 ivel = ld3(vel, id), iomega = ld3(omega, id);
    d3 f = ld3(force, id), mo = ld3(moment, id);
    for (uint32_t p = 0; p < nplane; ++p) {
        const xdem_device_plane_info& pl = planes[p];
        const xdem_device_contact_property& cp = cpp[p];
        d3 Fn = mk3(0, 0, 0), Ft = mk3(0, 0, 0), M = mk3(0, 0, 0);
        d3 dp = ipos3 - of(pl.xw);
        d3 unit = mk3(0, 0, 0);
        d3 wp = mk3(dot(dp, of(pl.u1)), dot(dp, of(pl.u2)), dot(dp, of(pl.uw)));
        double cdist = plane_detect(pl, ipos3, wp, unit, r);
        if (cdist > 0) {
            double rcon = r - 0.5 * cdist;
            d3 dv = k.rolling ? -(ivel + cross(iomega, r * unit)) : -ivel;
            ForceConst c = get_constant<TCM>(r, 0.0, m, 0.0, cp);
            force_law<TCM>(c, k.dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, M);
        }
        f += Fn + Ft;
        mo += M;
    }
    st3(force, id, f);
    st3(moment, id, mo);
}

// makeTFM_1..3 / toLocal / toGlobal — mphysics_cuda_impl.cuh:624-650
__device__ inline d3 tfm1(const xdem_double4& ep) {
    return mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w),
               2.0 * (ep.y * ep.w + ep.x * ep.z));
}
__device__ inline d3 tfm2(const xdem_double4& ep) {
    return mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5),
               2.0 * (ep.z * ep.w - ep.x * ep.y));
}
__device__ inline d3 tfm3(const xdem_double4& ep) {
    return mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y),
               2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
}
__device__ inline d3 to_local(d3 A1, d3 A2, d3 A3, d3 v) {
    return mk3(A1.x * v.x + A2.x * v.y + A3.x * v.z, A1.y * v.x + A2.y * v.y + A3.y * v.z,
               A1.z * v.x + A2.z * v.y + A3.z * v.z);
}
__device__ inline d3 to_global(d3 A1, d3 A2, d3 A3, d3 v) {
    return mk3(dot(A1, v), dot(A2, v), dot(A3, v));
}

// particle_cylinder_contact_detection — mphysics_cuda_impl.cuh:652-718 (the
// reference returns float: the penetration is rounded to FP32).
__device__ float cylinder_detect(const xdem_device_cylinder_info& cy, d4 pt, d3& u, d3& cp) {
    double dist = -1.0;
    d3 pbase = of(cy.pbase), ptop = of(cy.ptop), origin = of(cy.origin);
    d3 ab = mk3(ptop.x - pbase.x, ptop.y - pbase.y, ptop.z - pbase.z);
    d3 p = mk3(pt.x, pt.y, pt.z);
    double t = dot(p - pbase, ab) / dot(ab, ab);
    d3 _cp = mk3(0.0, 0.0, 0.0);
    if (t >= 0 && t <= 1) {
        _cp = pbase + t * ab;
        dist = length(p - _cp);
        u = (_cp - p) / dist;
        cp = _cp - cy.rbase * u;
        return cy.rtop + pt.w - dist;
    } else {
        _cp = pbase + t * ab;
        dist = length(p - _cp);
        if (dist < cy.rbase) {
            d3 OtoCp = origin - _cp;
            double OtoCp_ = length(OtoCp);
            u = OtoCp / OtoCp_;
            cp = _cp - cy.rbase * u;
            return cy.len * 0.5 + pt.w - OtoCp_;
        }
        d3 A_1 = tfm1(cy.ep), A_2 = tfm2(cy.ep), A_3 = tfm3(cy.ep);
        d3 _at = p - ptop;
        d3 at = to_local(A_1, A_2, A_3, _at);
        cp = ptop;
        if (fabs(at.y) > cy.len) {
            _at = p - pbase;
            at = to_local(A_1, A_2, A_3, _at);
            cp = pbase;
        }
        double pi = atan(at.x / at.z);
        if (pi < 0 && at.z < 0)
            _cp.x = cy.rbase * sin(-pi);
        else if (pi > 0 && at.x < 0 && at.z < 0)
            _cp.x = cy.rbase * sin(-pi);
        else
            _cp.x = cy.rbase * sin(pi);
        _cp.z = cy.rbase * cos(pi);
        if (at.z < 0 && _cp.z > 0)
            _cp.z = -_cp.z;
        else if (at.z > 0 && _cp.z < 0)
            _cp.z = -_cp.z;
        _cp.y = 0.;
        cp = cp + to_global(A_1, A_2, A_3, _cp);
        d3 disVec = cp - p;
        dist = length(disVec);
        u = disVec / dist;
        if (dist < pt.w) return pt.w - dist;
    }
    return -1.0;
}

// K10 cylinder_hertzian_contact_force_kernel<TCM> — :720-774 (no *mf race:
// every particle writes its own reaction).
template <int TCM>
__global__ __launch_bounds__(512) void k_cylinder(KParams k, const xdem_device_cylinder_info* cyp,
                                                  const double* pos, const double* vel,
                                                  const double* omega, double* force,
                                                  double* moment,
                                                  const xdem_device_contact_property* cpp,
                                                  const double* mass, const xdem_double3* mpos,
                                                  xdem_double3* mf, xdem_double3* mm, uint32_t np) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const xdem_device_cylinder_info cy = *cyp;
    double im = mass[id];
    d4 ipos = ld4(pos, id);
    d3 ivel = ld3(vel, id), iomega = ld3(omega, id);
    d3 unit = mk3(0, 0, 0), cpt = mk3(0, 0, 0);
    d3 mp = mpos ? of(*mpos) : mk3(0, 0, 0);
    double cdist = cylinder_detect(cy, ipos, unit, cpt);
    d3 si = cpt - mp;
    d3 cy2cp = cpt - of(cy.origin);
    d3 Ft = mk3(0, 0, 0), Fn = mk3(0, 0, 0), M = mk3(0, 0, 0);
    if (cdist > 0) {
        double rcon = ipos.w - 0.5 * cdist;
        d3 dv = of(cy.vel) + cross(of(cy.omega), cy2cp) - (ivel + cross(iomega, ipos.w * unit));
        ForceConst c = get_constant<TCM>(ipos.w, 0, im, 0, *cpp);
        force_law<TCM>(c, k.dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, M);
    }
    add3(force, id, Fn + Ft);
    add3(moment, id, M);
    if (mf) {
        d3 r = -(Fn);
        mf[id] = xdem_double3{r.x, r.y, r.z};
    }
    if (mm) {
        d3 r = cross(si, -Fn);
        mm[id] = xdem_double3{r.x, r.y, r.z};
    }
}

// Rotating drum (inner shell + optional end caps).  No reference kernel: the
// per-contact evaluation is that of cylinder_hertzian_contact_force_kernel
// (:720-774) — getConstant with the wall side empty, ir = 0, dv = wall point
// velocity - (v + w x (r n)) — applied to the inner surfaces.  The shell
// contact comes first, then the cap contact; their sum is added once.
template <int TCM>
__device__ inline void drum_one(const KParams& k, const xdem_drum_info& dr, double r, double im,
                                d3 iomega, d3 ivel, d3 unit, d3 cpt, double cdist,
                                const xdem_device_contact_property& cp, d3& F, d3& M, d3& R,
                                d3& RM) {
    d3 o = of(dr.origin);
    d3 wv = dr.omega * cross(of(dr.axis), cpt - o);
    d3 dv = wv - (ivel + cross(iomega, r * unit));
    double rcon = r - 0.5 * cdist;
    d3 Ft = mk3(0, 0, 0), Fn = mk3(0, 0, 0), Mc = mk3(0, 0, 0);
    ForceConst c = get_constant<TCM>(r, 0, im, 0, cp);
    force_law<TCM>(c, k.dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, Mc);
    d3 f = Fn + Ft;
    F += f;
    M += Mc;
    R += -f;
    RM += cross(cpt - o, -f);
}

template <int TCM>
__global__ __launch_bounds__(256) void k_drum(KParams k, const xdem_drum_info* drp,
                                              const double* pos, const double* vel,
                                              const double* omega, double* force, double* moment,
                                              const xdem_device_contact_property* cpp,
                                              const double* mass, xdem_double3* mf,
                                              xdem_double3* mm, uint32_t np) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    const xdem_drum_info dr = *drp;
    d4 ipos = ld4(pos, id);
    double r = ipos.w;
    d3 ax = of(dr.axis);
    d3 d = mk3(ipos.x, ipos.y, ipos.z) - of(dr.origin);
    double a = dot(d, ax);
    d3 q = d - a * ax;
    double rho = length(q);
    d3 F = mk3(0, 0, 0), M = mk3(0, 0, 0), R = mk3(0, 0, 0), RM = mk3(0, 0, 0);
    if (fabs(a) <= dr.half_length) {
        double im = mass[id];
        d3 ivel = ld3(vel, id), iomega = ld3(omega, id);
        double cs = r + rho - dr.radius;
        if (cs > 0 && rho > 0) {
            d3 u = q / rho;
            d3 cpt = of(dr.origin) + a * ax + dr.radius * u;
            drum_one<TCM>(k, dr, r, im, iomega, ivel, u, cpt, cs, *cpp, F, M, R, RM);
        }
        double cc = r - (dr.half_length - fabs(a));
        if (dr.caps && cc > 0 && rho < dr.radius) {
            double sg = a >= 0 ? 1.0 : -1.0;
            d3 u = sg * ax;
            d3 cpt = of(dr.origin) + (sg * dr.half_length) * ax + q;
            drum_one<TCM>(k, dr, r, im, iomega, ivel, u, cpt, cc, *cpp, F, M, R, RM);
        }
    }
    add3(force, id, F);
    add3(moment, id, M);
    if (mf) mf[id] = xdem_double3{R.x, R.y, R.z};
    if (mm) mm[id] = xdem_double3{RM.x, RM.y, RM.z};
}

// One lane stores a struct passed by value (a moved plane or polygon-object
// pose): no host staging buffer, so consecutive steps never race on it.
template <class T>
__global__ void k_put(T* dst, T v) {
    *dst = v;
}

// closestPtPointTriangle — mphysics_cuda_impl.cuh:837-899 (Ericson 5.1.5)
__device__ d3 closest_pt_triangle(const xdem_device_polygon_info& t, d3 p) {
    d3 a = of(t.P), b = of(t.Q), c = of(t.R);
    d3 ab = b - a, ac = c - a, ap = p - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    d3 bp = p - b;
    double d3_ = dot(ab, bp), d4_ = dot(ac, bp);
    if (d3_ >= 0.0 && d4_ <= d3_) return b;
    double vc = d1 * d4_ - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) {
        double v = d1 / (d1 - d3_);
        return a + v * ab;
    }
    d3 cp = p - c;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        return a + w * ac;
    }
    double va = d3_ * d6 - d5 * d4_;
    if (va <= 0.0 && (d4_ - d3_) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4_ - d3_) / ((d4_ - d3_) + (d5 - d6));
        return b + w * (c - b);
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom;
    double w = vc * denom;
    return a + v * ab + w * ac;
}

// K11 particle_polygonObject_collision_kernel<TCM> — :923-1021.  Reaction
// contributions per (object, particle) go to `react` (2*nobj*np double3) when
// non-null and are reduced in a fixed order afterwards (no racing +=).
template <int TCM>
__global__ __launch_bounds__(512) void k_polygon(
    KParams k, const xdem_device_polygon_info* dpi, const xdem_device_polygon_mass_info* dpmi,
    const double* pos, const double* vel, const double* omega, double* force, double* moment,
    const double* mass, const uint32_t* sorted_index, const uint32_t* cstart,
    const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np, uint32_t nobj,
    d3* react) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= np) return;
    double im = mass[id];
    d4 ip4 = ld4(pos, id);
    d3 ipos = mk3(ip4.x, ip4.y, ip4.z);
    d3 ivel = ld3(vel, id), iomega = ld3(omega, id);
    d3 unit = mk3(0, 0, 0);
    int gx, gy, gz;
    grid_pos(k, ipos.x, ipos.y, ipos.z, gx, gy, gz);
    double ir = ip4.w;
    d3 M = mk3(0, 0, 0), Fn = mk3(0, 0, 0), Ft = mk3(0, 0, 0);
    d3 sum_force = mk3(0, 0, 0), sum_moment = mk3(0, 0, 0);
    if (react)
        for (uint32_t o = 0; o < nobj; ++o) {
            react[(size_t)(2 * o) * np + id] = mk3(0, 0, 0);
            react[(size_t)(2 * o + 1) * np + id] = mk3(0, 0, 0);
        }
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t h = grid_hash(k, gx + x, gy + y, gz + z);
                uint32_t s = cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = cend[h];
                for (uint32_t j = s; j < e; j++) {
                    uint32_t kk = sorted_index[j];
                    if (kk < np) continue;
                    kk -= np;
                    const xdem_device_polygon_info tri = dpi[kk];
                    uint32_t pidx = (uint32_t)tri.id;
                    const xdem_device_polygon_mass_info& pmi = dpmi[pidx];
                    d3 cpt = closest_pt_triangle(tri, ipos);
                    d3 po2cp = cpt - of(pmi.origin);
                    d3 distVec = ipos - cpt;
                    double dist = length(distVec);
                    double cdist = ir - dist;
                    Fn = mk3(0.0, 0.0, 0.0);
                    if (cdist > 0) {
                        d3 qp = of(tri.Q) - of(tri.P);
                        d3 rp = of(tri.R) - of(tri.P);
                        double rcon = ir - 0.5 * cdist;
                        unit = -cross(qp, rp);
                        unit = unit / length(unit);
                        d3 dv = k.rolling ? of(pmi.vel) + cross(of(pmi.omega), po2cp) -
                                                (ivel + cross(iomega, ir * unit))
                                          : of(pmi.vel) - ivel;
                        ForceConst c = get_constant<TCM>(ir, 0, im, 0, cp[pidx]);
                        force_law<TCM>(c, k.dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, M);
                        sum_force += Fn + Ft;
                        sum_moment += M;
                        if (react && pidx < nobj) {
                            react[(size_t)(2 * pidx) * np + id] += -(Fn + Ft);
                            react[(size_t)(2 * pidx + 1) * np + id] += -cross(po2cp, Fn + Ft);
                        }
                    }
                }
            }
    add3(force, id, sum_force);
    add3(moment, id, sum_moment);
}

// Ordered block reduction of `rows` arrays of n d3 each into out[rows] (one
// block per row; fixed tree order, so the result is deterministic).
__global__ __launch_bounds__(512) void k_rows_sum(const d3* in, uint32_t n, d3* out) {
    __shared__ d3 s[512];
    const d3* row = in + (size_t)blockIdx.x * n;
    d3 acc = mk3(0, 0, 0);
    for (uint32_t i = threadIdx.x; i < n; i += 512) acc += row[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = 256; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// First pass of the row sums for long rows: block (b, row) sums the fixed
// chunk [b chunk, (b + 1) chunk) of its row in a fixed tree (deterministic).
constexpr uint32_t ROW_PARTS = 256;
__global__ __launch_bounds__(512) void k_rows_partial(const d3* in, uint32_t n, uint32_t chunk,
                                                      d3* partial) {
    __shared__ d3 s[512];
    const d3* row = in + (size_t)blockIdx.y * n;
    const uint64_t lo = (uint64_t)blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < n ? lo + chunk : n;
    d3 acc = mk3(0, 0, 0);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += 512) acc += row[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = 256; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * ROW_PARTS + blockIdx.x] = s[0];
}

// Largest polygon-object id over ntri triangles (objects = that + 1).
__global__ void k_max_object_id(const xdem_device_polygon_info* dpi, uint32_t ntri,
                                uint32_t* out) {
    uint32_t m = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntri;
         t += gridDim.x * blockDim.x)
        m = max(m, (uint32_t)dpi[t].id);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

__global__ void k_store_reactions(const d3* sums, xdem_device_polygon_mass_info* dpmi,
                                  uint32_t nobj) {
    uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nobj) return;
    d3 f = sums[2 * o], m = sums[2 * o + 1];
    dpmi[o].force = xdem_double3{f.x, f.y, f.z};
    dpmi[o].moment = xdem_double3{m.x, m.y, m.z};
}

// K13 reduce6<double3, 512> — mphysics_cuda_impl.cuh:776-835, launched as in
// reductionD3 (mphysics_cuda_def.cu:223-242).  Lanes beyond n count as zero.
__global__ __launch_bounds__(512) void k_reduce6(const d3* in, d3* out, uint32_t n,
                                                 uint32_t nthreads) {
    __shared__ d3 sdata[512];
    const uint32_t B = 512;
    uint32_t tid = threadIdx.x;
    uint64_t i = (uint64_t)blockIdx.x * (B * 2) + tid;
    uint64_t gridSize = (uint64_t)B * 2 * gridDim.x;
    d3 mySum = mk3(0.0, 0.0, 0.0);
    if (tid < nthreads) {
        while (i < n) {
            mySum += in[i];
            if (i + B < n) mySum += in[i + B];
            i += gridSize;
        }
    }
    sdata[tid] = mySum;
    __syncthreads();
    for (uint32_t s = 256; s >= 1; s >>= 1) {
        if (tid < s) sdata[tid] = mySum = mySum + sdata[tid + s];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = mySum;
}

// K12 updatePolygonObjectData_kernel — mphysics_cuda_impl.cuh:1898-1956
__global__ __launch_bounds__(512) void k_update_polygon(const xdem_device_polygon_mass_info* dpmi,
                                                        const double* vList, double* sphere,
                                                        xdem_device_polygon_info* dpi,
                                                        uint32_t ntri) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ntri) return;
    int mid = dpi[id].id;
    d3 pos = of(dpmi[mid].origin);
    xdem_double4 ep = dpmi[mid].ep;
    const double* s = vList + 9 * (size_t)id;
    d3 r0 = tfm1(ep), r1 = tfm2(ep), r2 = tfm3(ep);
    auto tg = [&](d3 v) {
        return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
                   r2.x * v.x + r2.y * v.y + r2.z * v.z);
    };
    d3 P = pos + tg(mk3(s[0], s[1], s[2]));
    d3 Q = pos + tg(mk3(s[3], s[4], s[5]));
    d3 R = pos + tg(mk3(s[6], s[7], s[8]));
    d3 V = Q - P, W = R - P;
    d3 N = cross(V, W);
    N = N / length(N);
    d3 M1 = 0.5 * (Q + P), M2 = 0.5 * (R + P);
    d3 D1 = cross(N, V), D2 = cross(N, W);
    double t = 0.0;
    if (fabs(D1.x * D2.y - D1.y * D2.x) > 1E-13)
        t = (D2.x * (M1.y - M2.y)) / (D1.x * D2.y - D1.y * D2.x) -
            (D2.y * (M1.x - M2.x)) / (D1.x * D2.y - D1.y * D2.x);
    else if (fabs(D1.x * D2.z - D1.z * D2.x) > 1E-13)
        t = (D2.x * (M1.z - M2.z)) / (D1.x * D2.z - D1.z * D2.x) -
            (D2.z * (M1.x - M2.x)) / (D1.x * D2.z - D1.z * D2.x);
    else if (fabs(D1.y * D2.z - D1.z * D2.y) > 1E-13)
        t = (D2.y * (M1.z - M2.z)) / (D1.y * D2.z - D1.z * D2.y) -
            (D2.z * (M1.y - M2.y)) / (D1.y * D2.z - D1.z * D2.y);
    d3 ctri = M1 + t * D1;
    sphere[4 * (size_t)id + 0] = ctri.x;
    sphere[4 * (size_t)id + 1] = ctri.y;
    sphere[4 * (size_t)id + 2] = ctri.z;
    dpi[id].P = xdem_double3{P.x, P.y, P.z};
    dpi[id].Q = xdem_double3{Q.x, Q.y, Q.z};
    dpi[id].R = xdem_double3{R.x, R.y, R.z};
}

int32_t need_params() {
    if (!params_set()) {
        set_error("parameters never set (xdem_set_symbolic_parameter)");
        return XDEM_ERR_STATE;
    }
    return XDEM_OK;
}

}  // namespace

// ------------------------------------------------------------ launchers
int32_t launch_vv_position(const KParams& k, double* pos, double* vel, double* acc, uint32_t np,
                           hipStream_t s) {
    if (!np) return XDEM_OK;
    hipLaunchKernelGGL(k_vv_position, dim3(div_up(np, 512)), dim3(512), 0, s, k, pos, vel, acc, np);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_vv_velocity(const KParams& k, double* vel, double* acc, double* omega,
                           double* alpha, double* force, double* moment, const double* mass,
                           const double* iner, uint32_t np, hipStream_t s) {
    if (!np) return XDEM_OK;
    hipLaunchKernelGGL(k_vv_velocity, dim3(div_up(np, 512)), dim3(512), 0, s, k, vel, acc, omega,
                       alpha, force, moment, mass, iner, np);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_hash(const KParams& k, uint32_t* hash, uint32_t* index, const double* pos,
                    uint32_t sid, uint32_t n, hipStream_t s) {
    if (!n) return XDEM_OK;
    hipLaunchKernelGGL(k_hash, dim3(div_up(n, 512)), dim3(512), 0, s, k, hash, index, pos, sid, n);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_p2p(int tcm, const KParams& k, const double* pos, const double* vel,
                   const double* omega, double* force, double* moment, const double* mass,
                   const uint32_t* sorted, const uint32_t* cstart, const uint32_t* cend,
                   const xdem_device_contact_property* cp, uint32_t np, uint32_t nsorted,
                   hipStream_t s, float4* pfs) {
    if (!np) return XDEM_OK;
    const dim3 g(div_up(nsorted ? nsorted : np, 256));
    // screening range: twice the grid's extent from the origin
    const double pf_range = 2.0 * k.cs * std::max(k.gx, std::max(k.gy, k.gz));
    const float pf_eps = (float)std::ldexp(pf_range, -20);
    if (nsorted) {
        if (pfs) {
            hipLaunchKernelGGL(k_sorted_screen, dim3(div_up(nsorted, 256)), dim3(256), 0, s, k, pos,
                               sorted, np, nsorted, pf_range, pfs);
            XDEM_LAUNCHED();
        }
        if (tcm == 0)
            hipLaunchKernelGGL((k_p2p<0, true>), g, dim3(256), 0, s, k, pos, vel, omega, force,
                               moment, mass, sorted, cstart, cend, cp, np, nsorted, pfs, pf_range,
                               pf_eps);
        else
            hipLaunchKernelGGL((k_p2p<1, true>), g, dim3(256), 0, s, k, pos, vel, omega, force,
                               moment, mass, sorted, cstart, cend, cp, np, nsorted, pfs, pf_range,
                               pf_eps);
    } else if (tcm == 0) {
        hipLaunchKernelGGL((k_p2p<0, false>), g, dim3(256), 0, s, k, pos, vel, omega, force,
                           moment, mass, sorted, cstart, cend, cp, np, 0u, nullptr, pf_range, pf_eps);
    } else {
        hipLaunchKernelGGL((k_p2p<1, false>), g, dim3(256), 0, s, k, pos, vel, omega, force,
                           moment, mass, sorted, cstart, cend, cp, np, 0u, nullptr, pf_range, pf_eps);
    }
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_planes(int tcm, const KParams& k, const xdem_device_plane_info* planes,
                      const xdem_device_contact_property* cps, uint32_t nplane, const double* pos,
                      const double* vel, const double* omega, double* force, double* moment,
                      const double* mass, uint32_t np, hipStream_t s) {
    if (!np || !nplane) return XDEM_OK;
    if (tcm == 0)
        hipLaunchKernelGGL(k_plane<0>, dim3(div_up(np, 256)), dim3(256), 0, s, k, planes, nplane,
                           pos, vel, omega, force, moment, cps, mass, np);
    else
        hipLaunchKernelGGL(k_plane<1>, dim3(div_up(np, 256)), dim3(256), 0, s, k, planes, nplane,
                           pos, vel, omega, force, moment, cps, mass, np);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_cylinder(int tcm, const KParams& k, const xdem_device_cylinder_info* cy,
                        const double* pos, const double* vel, const double* omega, double* force,
                        double* moment, const double* mass, uint32_t np,
                        const xdem_device_contact_property* cp, const xdem_double3* mpos,
                        xdem_double3* mf, xdem_double3* mm, hipStream_t s) {
    if (!np) return XDEM_OK;
    if (tcm == 0)
        hipLaunchKernelGGL(k_cylinder<0>, dim3(div_up(np, 512)), dim3(512), 0, s, k, cy, pos, vel,
                           omega, force, moment, cp, mass, mpos, mf, mm, np);
    else
        hipLaunchKernelGGL(k_cylinder<1>, dim3(div_up(np, 512)), dim3(512), 0, s, k, cy, pos, vel,
                           omega, force, moment, cp, mass, mpos, mf, mm, np);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_drum(int tcm, const KParams& k, const xdem_drum_info* dr, const double* pos,
                    const double* vel, const double* omega, double* force, double* moment,
                    const double* mass, uint32_t np, const xdem_device_contact_property* cp,
                    xdem_double3* mf, xdem_double3* mm, hipStream_t s) {
    if (!np) return XDEM_OK;
    if (tcm == 0)
        hipLaunchKernelGGL(k_drum<0>, dim3(div_up(np, 256)), dim3(256), 0, s, k, dr, pos, vel,
                           omega, force, moment, cp, mass, mf, mm, np);
    else
        hipLaunchKernelGGL(k_drum<1>, dim3(div_up(np, 256)), dim3(256), 0, s, k, dr, pos, vel,
                           omega, force, moment, cp, mass, mf, mm, np);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_put_plane(xdem_device_plane_info* dst, const xdem_device_plane_info& v,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_put<xdem_device_plane_info>, dim3(1), dim3(1), 0, s, dst, v);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_put_object(xdem_device_polygon_mass_info* dst,
                          const xdem_device_polygon_mass_info& v, hipStream_t s) {
    hipLaunchKernelGGL(k_put<xdem_device_polygon_mass_info>, dim3(1), dim3(1), 0, s, dst, v);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t launch_polygon(int tcm, const KParams& k, const xdem_device_polygon_info* dpi,
                       xdem_device_polygon_mass_info* dpmi, uint32_t nobj, const double* pos,
                       const double* vel, const double* omega, double* force, double* moment,
                       const double* mass, const uint32_t* sorted, const uint32_t* cstart,
                       const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np,
                       hipStream_t s) {
    if (!np) return XDEM_OK;
    d3* react = nullptr;
    d3* sums = nullptr;
    d3* partial = nullptr;
    if (nobj) {
        XDEM_TRY(workspace(1).reserve(sizeof(d3) * (2 * (size_t)nobj * np + 2 * nobj +
                                                    2 * (size_t)nobj * ROW_PARTS)));
        react = static_cast<d3*>(workspace(1).ptr);
        sums = react + 2 * (size_t)nobj * np;
        partial = sums + 2 * nobj;
    }
    if (tcm == 0)
        hipLaunchKernelGGL(k_polygon<0>, dim3(div_up(np, 512)), dim3(512), 0, s, k, dpi, dpmi, pos,
                           vel, omega, force, moment, mass, sorted, cstart, cend, cp, np, nobj,
                           react);
    else
        hipLaunchKernelGGL(k_polygon<1>, dim3(div_up(np, 512)), dim3(512), 0, s, k, dpi, dpmi, pos,
                           vel, omega, force, moment, mass, sorted, cstart, cend, cp, np, nobj,
                           react);
    XDEM_LAUNCHED();
    if (nobj) {
        const uint32_t chunk = div_up(np, ROW_PARTS);
        hipLaunchKernelGGL(k_rows_partial, dim3(ROW_PARTS, 2 * nobj), dim3(512), 0, s, react, np,
                           chunk, partial);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(k_rows_sum, dim3(2 * nobj), dim3(512), 0, s, partial, ROW_PARTS, sums);
        XDEM_LAUNCHED();
        hipLaunchKernelGGL(k_store_reactions, dim3(div_up(nobj, 64)), dim3(64), 0, s, sums, dpmi,
                           nobj);
        XDEM_LAUNCHED();
    }
    return XDEM_OK;
}

int32_t launch_update_polygon(const xdem_device_polygon_mass_info* dpmi, const double* vList,
                              double* sphere, xdem_device_polygon_info* dpi, uint32_t ntri,
                              hipStream_t s) {
    if (!ntri) return XDEM_OK;
    hipLaunchKernelGGL(k_update_polygon, dim3(div_up(ntri, 512)), dim3(512), 0, s, dpmi, vList,
                       sphere, dpi, ntri);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

int32_t reduction_d3(const d3* in, uint32_t n, d3* out, hipStream_t s) {
    *out = mk3(0, 0, 0);
    if (!n) return XDEM_OK;
    uint32_t nthreads = n < 512 ? n : 512;
    uint32_t nblocks = div_up(n, nthreads);
    XDEM_TRY(workspace(2).reserve(sizeof(d3) * nblocks));
    d3* dout = static_cast<d3*>(workspace(2).ptr);
    hipLaunchKernelGGL(k_reduce6, dim3(nblocks), dim3(512), 0, s, in, dout, n, nthreads);
    XDEM_LAUNCHED();
    std::vector<d3> h(nblocks);
    XDEM_HIP(hipMemcpyAsync(h.data(), dout, sizeof(d3) * nblocks, hipMemcpyDeviceToHost, s));
    XDEM_HIP(hipStreamSynchronize(s));
    d3 rt = mk3(0.0, 0.0, 0.0);
    for (uint32_t i = 0; i < nblocks; ++i) {
        rt.x += h[i].x;
        rt.y += h[i].y;
        rt.z += h[i].z;
    }
    *out = rt;
    return XDEM_OK;
}

int32_t launch_pairs(const KParams& k, const double* pos, const uint32_t* sorted,
                     const uint32_t* cstart, const uint32_t* cend, uint32_t np, uint32_t max_per,
                     uint32_t* counts, uint32_t* partners, hipStream_t s) {
    if (!np) return XDEM_OK;
    hipLaunchKernelGGL(k_pairs, dim3(div_up(np, 256)), dim3(256), 0, s, k, pos, sorted, cstart,
                       cend, np, max_per, counts, partners);
    XDEM_LAUNCHED();
    return XDEM_OK;
}

}  // namespace xdem

using namespace xdem;

// ================================================================ Set A
extern "C" {

int32_t xdem_vv_update_position_f64(double* pos, double* vel, double* acc, uint32_t np) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(np == 0 || (pos && vel && acc), "null array");
    return launch_vv_position(params_from(host_params()), pos, vel, acc, np, stream());
}

int32_t xdem_vv_update_velocity_f64(double* vel, double* acc, double* omega, double* alpha,
                                    double* force, double* moment, double* mass, double* iner,
                                    uint32_t np) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(np == 0 || (vel && acc && omega && alpha && force && moment && mass && iner),
                 "null array");
    return launch_vv_velocity(params_from(host_params()), vel, acc, omega, alpha, force, moment,
                              mass, iner, np, stream());
}

int32_t xdem_calculate_hash_and_index_f64(uint32_t* hash, uint32_t* index, const double* pos,
                                          uint32_t np) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(np == 0 || (hash && index && pos), "null array");
    return launch_hash(params_from(host_params()), hash, index, pos, 0, np, stream());
}

int32_t xdem_calculate_hash_and_index_polygon_sphere_f64(uint32_t* hash, uint32_t* index,
                                                         uint32_t sid, uint32_t nsphere,
                                                         const double* sphere) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(nsphere == 0 || (hash && index && sphere), "null array");
    return launch_hash(params_from(host_params()), hash, index, sphere, sid, nsphere, stream());
}

int32_t xdem_calculate_p2p_f64(int32_t tcm, const double* pos, const double* vel,
                               const double* omega, double* force, double* moment,
                               const double* mass, const uint32_t* sorted_index,
                               const uint32_t* cstart, const uint32_t* cend,
                               const xdem_device_contact_property* cp, uint32_t np) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (pos && vel && omega && force && moment && mass && sorted_index &&
                             cstart && cend && cp),
                 "null array");
    return launch_p2p(tcm, params_from(host_params()), pos, vel, omega, force, moment, mass,
                      sorted_index, cstart, cend, cp, np, 0u, stream());
}

int32_t xdem_plane_contact_force_f64(int32_t tcm, const xdem_device_plane_info* plane,
                                     const double* pos, const double* vel, const double* omega,
                                     double* force, double* moment, const double* mass,
                                     uint32_t np, const xdem_device_contact_property* cp) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (plane && pos && vel && omega && force && moment && mass && cp),
                 "null array");
    // One plane with one contact property.
    return launch_planes(tcm, params_from(host_params()), plane, cp, 1, pos, vel, omega, force,
                         moment, mass, np, stream());
}

int32_t xdem_cube_contact_force_f64(int32_t tcm, const xdem_device_plane_info* planes6,
                                    const double* pos, const double* vel, const double* omega,
                                    double* force, double* moment, const double* mass,
                                    uint32_t np, const xdem_device_contact_property* cp) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (planes6 && pos && vel && omega && force && moment && mass && cp),
                 "null array");
    // mphysics_cuda_def.cu:158-179: six launches with the same property.
    KParams k = params_from(host_params());
    for (int i = 0; i < 6; ++i)
        XDEM_TRY(launch_planes(tcm, k, planes6 + i, cp, 1, pos, vel, omega, force, moment, mass, np,
                               stream()));
    return XDEM_OK;
}

int32_t xdem_particle_polygon_object_collision_f64(
    int32_t tcm, const xdem_device_polygon_info* dpi, const double* dsph,
    xdem_device_polygon_mass_info* dpmi, uint32_t nobject, uint32_t ntriangle, const double* pos,
    const double* vel, const double* omega, double* force, double* moment, const double* mass,
    const uint32_t* sorted_index, const uint32_t* cstart, const uint32_t* cend,
    const xdem_device_contact_property* cp, uint32_t np) {
    (void)dsph;
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (dpi && dpmi && pos && vel && omega && force && moment && mass &&
                             sorted_index && cstart && cend && cp),
                 "null array");
    XDEM_REQUIRE(host_params().np == np, "np (%u) must equal the parameter block np (%u): "
                 "triangle entries are the sorted indices >= np", np, host_params().np);
    if (nobject == XDEM_NOBJECT_AUTO) {
        // the objects the triangles name: 1 + the largest dpi[t].id
        const uint32_t ntri = ntriangle ? ntriangle : host_params().nsphere;
        nobject = 0;
        if (ntri && np) {
            XDEM_REQUIRE(dpi, "null dpi");
            hipStream_t s = stream();
            XDEM_TRY(workspace(3).reserve(sizeof(uint32_t)));
            uint32_t* d = static_cast<uint32_t*>(workspace(3).ptr);
            XDEM_TRY(fill_u32(d, 0u, 1, s));
            hipLaunchKernelGGL(k_max_object_id, dim3(std::min<uint32_t>(div_up(ntri, 256), 1024)),
                               dim3(256), 0, s, dpi, ntri, d);
            XDEM_LAUNCHED();
            uint32_t mx = 0;
            XDEM_HIP(hipMemcpyAsync(&mx, d, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            XDEM_HIP(hipStreamSynchronize(s));
            XDEM_REQUIRE(mx < (1u << 16), "polygon-object id %u out of range (ids must be < 65536)",
                         mx);
            nobject = mx + 1;
        }
    }
    return launch_polygon(tcm, params_from(host_params()), dpi, dpmi, nobject, pos, vel, omega,
                          force, moment, mass, sorted_index, cstart, cend, cp, np, stream());
}

int32_t xdem_cylinder_hertzian_contact_force_f64(
    int32_t tcm, const xdem_device_cylinder_info* cyl, const double* pos, const double* vel,
    const double* omega, double* force, double* moment, const double* mass, uint32_t np,
    const xdem_device_contact_property* cp, const xdem_double3* mpos, xdem_double3* mf,
    xdem_double3* mm, xdem_double3* out_mf, xdem_double3* out_mm) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (cyl && pos && vel && omega && force && moment && mass && cp),
                 "null array");
    hipStream_t s = stream();
    XDEM_TRY(launch_cylinder(tcm, params_from(host_params()), cyl, pos, vel, omega, force, moment,
                             mass, np, cp, mpos, mf, mm, s));
    if (out_mf && mf) {
        d3 r;
        XDEM_TRY(reduction_d3(reinterpret_cast<const d3*>(mf), np, &r, s));
        *out_mf = xdem_double3{r.x, r.y, r.z};
    }
    if (out_mm && mm) {
        d3 r;
        XDEM_TRY(reduction_d3(reinterpret_cast<const d3*>(mm), np, &r, s));
        *out_mm = xdem_double3{r.x, r.y, r.z};
    }
    return XDEM_OK;
}

int32_t xdem_drum_contact_force_f64(int32_t tcm, const xdem_drum_info* drum, const double* pos,
                                    const double* vel, const double* omega, double* force,
                                    double* moment, const double* mass, uint32_t np,
                                    const xdem_device_contact_property* cp, xdem_double3* mf,
                                    xdem_double3* mm, xdem_double3* out_mf, xdem_double3* out_mm) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(tcm == 0 || tcm == 1, "tcm must be 0 (HMCM) or 1 (DHS)");
    XDEM_REQUIRE(np == 0 || (drum && pos && vel && omega && force && moment && mass && cp),
                 "null array");
    hipStream_t s = stream();
    XDEM_TRY(launch_drum(tcm, params_from(host_params()), drum, pos, vel, omega, force, moment,
                         mass, np, cp, mf, mm, s));
    if (out_mf && mf) {
        d3 r;
        XDEM_TRY(reduction_d3(reinterpret_cast<const d3*>(mf), np, &r, s));
        *out_mf = xdem_double3{r.x, r.y, r.z};
    }
    if (out_mm && mm) {
        d3 r;
        XDEM_TRY(reduction_d3(reinterpret_cast<const d3*>(mm), np, &r, s));
        *out_mm = xdem_double3{r.x, r.y, r.z};
    }
    return XDEM_OK;
}

int32_t xdem_reduction_d3(const xdem_double3* in, uint32_t n, xdem_double3* out) {
    XDEM_REQUIRE(out && (n == 0 || in), "null array");
    d3 r;
    XDEM_TRY(reduction_d3(reinterpret_cast<const d3*>(in), n, &r, stream()));
    *out = xdem_double3{r.x, r.y, r.z};
    return XDEM_OK;
}

int32_t xdem_update_polygon_object_data_f64(const xdem_device_polygon_mass_info* dpmi,
                                            const double* vList, double* sphere,
                                            xdem_device_polygon_info* dpi, uint32_t ntriangle) {
    XDEM_REQUIRE(ntriangle == 0 || (dpmi && vList && sphere && dpi), "null array");
    return launch_update_polygon(dpmi, vList, sphere, dpi, ntriangle, stream());
}

int32_t xdem_contact_pairs_f64(const double* pos, const uint32_t* sorted_index,
                               const uint32_t* cstart, const uint32_t* cend, uint32_t np,
                               uint32_t max_per_particle, uint32_t* counts, uint32_t* partners) {
    XDEM_TRY(need_params());
    XDEM_REQUIRE(np == 0 || (pos && sorted_index && cstart && cend && counts && partners),
                 "null array");
    XDEM_REQUIRE(max_per_particle > 0, "max_per_particle must be > 0");
    return launch_pairs(params_from(host_params()), pos, sorted_index, cstart, cend, np,
                        max_per_particle, counts, partners, stream());
}

}  // extern "C"
