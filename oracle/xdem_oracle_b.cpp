// xdem_oracle_b.cpp — CPU FP64 restatement of the reference's older DEM
// entry points (include/cu_dem_dec.cuh, source/cu_dem_def.cu,
// include/cu_dem_impl.cuh).
//
// TEST INFRASTRUCTURE: the checker for the xdem_b_* entry points of
// libxdem.so.  Nothing in the product links or calls this file.
//
// Launch bounds.  A reference kernel launched for n items with
// computeGridSize(n, 256) (cu_dem_def.cu:19-23) runs T(n) = numBlocks *
// numThreads >= n threads and each thread returns when its id reaches the
// kernel's bound (cte.np, or cte.np + cte.nsp); so the ids processed are
// those below min(T(n), bound).  Restated as such.
//
// Vector helpers follow cu_dem_impl.cuh:10-71: operator/ multiplies by the
// reciprocal (:68-71), length and dot sum left to right, sign(0) = +1 (:10).
// Departures (defects not reproduced): np == 0 is a no-op (the reference
// divides by zero in computeGridSize); reduce6 lanes beyond the particle count
// read zero (uninitialised shared memory in the reference when np < 256).
//
// Pins (tests/test_oracle_setb.py): hash/sort/cell ranges = the live-set
// oracle and the stable-sort definition; Verlet = exact parabola and the
// two-half-kick identity; p2p = head-on Hertz closed form, sliding moment
// closed form, cohesion closed form (force step tu = coh / mu along the
// centre line); cube = sphere resting on a face (kn d^1.5), nothing outside
// the face rectangle; shape = face contact closed form with the triangle
// normal, edge contact along the centre line, one contact per particle,
// body force = the sum of the particle forces; polygon updates = rotation
// closed forms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "xdem_oracle.h"

namespace {

struct V3 {
    double x, y, z;
};
inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
inline V3 ld(const double* p) { return V3{p[0], p[1], p[2]}; }
inline V3 of(const xdem_double3& d) { return V3{d.x, d.y, d.z}; }
inline void st(double* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double b) {  // cu_dem_impl.cuh:68-71
    double inv = 1 / b;
    return V3{a.x * inv, a.y * inv, a.z * inv};
}
inline void operator+=(V3& a, V3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline int sgn(double a) { return a < 0 ? -1 : 1; }  // cu_dem_impl.cuh:10-13
inline uint32_t umul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

// computeGridSize(n, 256) thread count (cu_dem_def.cu:13-23)
uint64_t threads_for(uint32_t n) {
    if (n == 0) return 0;
    uint32_t t = std::min<uint32_t>(256, n);
    return (uint64_t)((n + t - 1) / t) * t;
}
uint32_t ids_below(uint32_t n, uint64_t bound) { return (uint32_t)std::min<uint64_t>(threads_for(n), bound); }

// calcGridPos / calcGridHash — cu_dem_impl.cuh:73-91
uint32_t hash_of(const xdem_b_device_parameters& p, int cx, int cy, int cz) {
    uint32_t x = (uint32_t)cx & (p.grid_size.x - 1);
    uint32_t y = (uint32_t)cy & (p.grid_size.y - 1);
    uint32_t z = (uint32_t)cz & (p.grid_size.z - 1);
    return umul24(umul24(z, p.grid_size.y), p.grid_size.x) + umul24(y, p.grid_size.x) + x;
}
void cell_of(const xdem_b_device_parameters& p, V3 q, int& cx, int& cy, int& cz) {
    cx = (int)std::floor((q.x - p.world_origin.x) / p.cell_size);
    cy = (int)std::floor((q.y - p.world_origin.y) / p.cell_size);
    cz = (int)std::floor((q.z - p.world_origin.z) / p.cell_size);
}

// calForce — cu_dem_impl.cuh:198-252
bool cal_force(const xdem_b_device_parameters& c, double ir, double jr, V3 ipos, V3 jpos, V3 ivel,
               V3 jvel, V3 iomega, V3 jomega, V3& force, V3& moment) {
    V3 rp = jpos - ipos;
    double dist = length(rp);
    double cd = (ir + jr) - dist;
    if (cd <= 0) {
        if (c.cohesive == 0) return false;
        double tu = c.cohesive / c.mu;
        V3 unit = rp / dist;
        V3 rv = jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit));
        double Fn = (-c.kn * std::pow(-cd, 1.5) + c.vn * dot(rv, unit));
        if (Fn < 0 && Fn > -tu) force += tu * unit;
        V3 e = rv - dot(rv, unit) * unit;
        double mag_e = length(e);
        if (mag_e) {
            V3 s_hat = e / mag_e;
            double ds = mag_e * c.dt;
            double Fs = std::fmin(c.ks * ds + c.vs * (dot(rv, s_hat)), c.mu * Fn);
            if (Fs < Fn * c.mu + c.cohesive) moment += cross(ir * unit, (Fn * c.mu + c.cohesive) * s_hat);
        }
        return false;
    }
    V3 unit = rp / dist;
    V3 rv = jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit));
    V3 single_force = (-c.kn * std::pow(cd, 1.5) + c.vn * dot(rv, unit)) * unit;
    V3 single_moment = mk(0, 0, 0);
    V3 e = rv - dot(rv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        V3 s_hat = e / mag_e;
        double ds = mag_e * c.dt;
        V3 shear = std::fmin(c.ks * ds + c.vs * (dot(rv, s_hat)), c.mu * length(single_force)) * s_hat;
        single_moment = cross(ir * unit, shear);
    }
    force += single_force;
    moment += single_moment;
    return true;
}

// ClosestPtPointTriangle with its region code (cu_dem_impl.cuh:380-441):
// 0 vertex, 1 edge, 2 face.
V3 closest_pt(V3 p, V3 a, V3 b, V3 c, int* wc) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        *wc = 0;
        return a;
    }
    V3 bp = p - b;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        *wc = 0;
        return b;
    }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        *wc = 1;
        double v = d1 / (d1 - d3);
        return a + v * ab;
    }
    V3 cp = p - c;
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
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        *wc = 1;
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + w * (c - b);
    }
    *wc = 2;
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom;
    double w = vc * denom;
    return a + v * ab + w * ac;
}

// The force of one triangle contact (the shared body of
// calculateTriangleParticleContactForce :454-526 and
// calculateTriangleLineOrEdgeParticleContactForce :528-589).  The moment is
// assigned only when the tangential velocity is non-zero.
void tri_force(const xdem_b_device_parameters& c, double kn, double vn, double ks, double vs,
               double mu, double r, double cd, V3 unit, V3 vel, V3 omega, V3 body_vel,
               V3& force, V3& moment) {
    V3 dv = body_vel - (vel + cross(omega, r * unit));
    force = (-kn * std::pow(cd, 1.5) + vn * dot(dv, unit)) * unit;
    V3 e = dv - dot(dv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        V3 s_hat = e / mag_e;
        double ds = mag_e * c.dt;
        V3 shear = std::fmin(ks * ds + vs * dot(dv, s_hat), mu * length(force)) * s_hat;
        moment = cross(r * unit, shear);
    }
}

struct ShapeIn {
    const xdem_b_device_parameters* p;
    const xdem_b_mass_info* mi;
    const xdem_b_polygon* polys;
    const uint32_t *id_set, *poly_start, *poly_end;
    double kn, vn, ks, vs, mu;
    const double *pos4, *vel4, *omega4;
    const uint32_t *sorted, *cstart, *cend;
};

// One pass of the 27-cell traversal over the shape vertices (k >= np) of
// particle id; `face` selects the face pass (:676-762, first face-region
// contact) or the line/edge pass (:591-674, first vertex/edge-region contact).
// Returns true with (F, M) of the contact taken.
bool shape_pass(const ShapeIn& S, uint32_t id, bool face, V3& F, V3& M) {
    const xdem_b_device_parameters& c = *S.p;
    V3 ipos = ld(S.pos4 + 4 * id), ivel = ld(S.vel4 + 4 * id), iomega = ld(S.omega4 + 4 * id);
    double ir = S.pos4[4 * id + 3];
    int gx, gy, gz;
    cell_of(c, ipos, gx, gy, gz);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t h = hash_of(c, gx + x, gy + y, gz + z);
                uint32_t s = S.cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = S.cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = S.sorted[j];
                    if (id == k || k < c.np) continue;
                    uint32_t vid = k - c.np;
                    for (uint32_t i = S.poly_start[vid]; i < S.poly_end[vid]; ++i) {
                        const xdem_b_polygon& po = S.polys[S.id_set[i]];
                        int wc = -1;
                        V3 cp = closest_pt(ipos, of(po.P), of(po.Q), of(po.R), &wc);
                        V3 sp = ipos - cp;
                        double dist = length(sp);
                        double cd = ir - dist;
                        if (cd <= 0) continue;
                        if (face && wc == 2) {
                            // the face pass returns at its first contact, so
                            // plane_contact_count is 0 here and the line
                            // intersection test (:494-500) never runs
                            V3 unit = -of(po.N) / length(of(po.N));
                            F = mk(0, 0, 0);
                            M = mk(0, 0, 0);
                            V3 f, m = mk(0, 0, 0);
                            tri_force(c, S.kn, S.vn, S.ks, S.vs, S.mu, ir, cd, unit, ivel, iomega,
                                      of(S.mi->vel), f, m);
                            F += f;
                            M = m;
                            return true;
                        }
                        if (!face && (wc == 0 || wc == 1)) {
                            V3 unit = -sp / dist;
                            V3 f, m = mk(0, 0, 0);
                            tri_force(c, S.kn, S.vn, S.ks, S.vs, S.mu, ir, cd, unit, ivel, iomega,
                                      of(S.mi->vel), f, m);
                            F = f;
                            M = m;
                            return true;
                        }
                    }
                }
            }
    return false;
}

// reduce6<double3, 256> over n items with computeGridSize(n, 256) blocks
// (cu_dem_impl.cuh:820-847; cu_dem_def.cu:214-225): per block the strided
// sums, then the halving tree; the host adds the block results in order.
void reduce6_256(const double* in3, uint32_t n, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (n == 0) return;
    const uint32_t B = 256;
    uint32_t nthreads = std::min(n, B);
    uint32_t nblocks = (n + nthreads - 1) / nthreads;
    uint64_t grid = (uint64_t)B * 2 * nblocks;
    std::vector<V3> sd(B);
    for (uint32_t b = 0; b < nblocks; ++b) {
        for (uint32_t t = 0; t < B; ++t) {
            V3 s = mk(0, 0, 0);
            if (t < nthreads) {
                uint64_t i = (uint64_t)b * (B * 2) + t;
                while (i < n) {
                    s += ld(in3 + 3 * i);
                    if (i + B < n) s += ld(in3 + 3 * (i + B));
                    i += grid;
                }
            }
            sd[t] = s;
        }
        for (uint32_t h = B / 2; h >= 1; h >>= 1)
            for (uint32_t t = 0; t < h; ++t) sd[t] = sd[t] + sd[t + h];
        out[0] += sd[0].x;
        out[1] += sd[0].y;
        out[2] += sd[0].z;
    }
}

}  // namespace

extern "C" {

// vv_update_position_kernel — cu_dem_impl.cuh:93-107
void orcb_vv_position(const xdem_b_device_parameters* p, double* pos4, const double* vel4,
                      const double* acc4, uint32_t np) {
    uint32_t n = ids_below(np, p->np);
    for (uint32_t i = 0; i < n; ++i) {
        V3 q = ld(pos4 + 4 * i), v = ld(vel4 + 4 * i), a = ld(acc4 + 4 * i);
        q += p->dt * v + p->half2dt * a;
        st(pos4 + 4 * i, q);
    }
}

// vv_update_velocity_kernel — cu_dem_impl.cuh:109-140 (two half kicks)
void orcb_vv_velocity(const xdem_b_device_parameters* p, double* vel4, double* acc4,
                      double* omega4, double* alpha4, const double* force3,
                      const double* moment3, uint32_t np) {
    uint32_t n = ids_below(np, p->np);
    for (uint32_t i = 0; i < n; ++i) {
        V3 v = ld(vel4 + 4 * i), a = ld(acc4 + 4 * i), av = ld(omega4 + 4 * i),
           aa = ld(alpha4 + 4 * i);
        double mass = acc4[4 * i + 3], inertia = alpha4[4 * i + 3];
        v += 0.5 * p->dt * a;
        av += 0.5 * p->dt * aa;
        a = (1 / mass) * ld(force3 + 3 * i);
        aa = (1 / inertia) * ld(moment3 + 3 * i);
        v += 0.5 * p->dt * a;
        av += 0.5 * p->dt * aa;
        st(vel4 + 4 * i, v);
        vel4[4 * i + 3] = 0;
        st(omega4 + 4 * i, av);
        omega4[4 * i + 3] = 0;
        st(acc4 + 4 * i, a);
        acc4[4 * i + 3] = mass;
        st(alpha4 + 4 * i, aa);
        alpha4[4 * i + 3] = inertia;
    }
}

// mergedata_kernel — cu_dem_impl.cuh:782-796
void orcb_mergedata(double* data4, const double* pos4, uint32_t np, const double* vertice3,
                    uint32_t snp) {
    uint32_t n = (uint32_t)std::min<uint64_t>(threads_for(np + snp), (uint64_t)np + snp);
    for (uint32_t i = 0; i < n; ++i) {
        if (i < np) {
            std::memcpy(data4 + 4 * i, pos4 + 4 * i, 32);
        } else {
            const double* v = vertice3 + 3 * (size_t)(i - np);
            data4[4 * i] = v[0];
            data4[4 * i + 1] = v[1];
            data4[4 * i + 2] = v[2];
            data4[4 * i + 3] = -1.0;
        }
    }
}

// cu_calculateHashAndIndex — cu_dem_def.cu:57-69 with calculateHashAndIndex_kernel
// (cu_dem_impl.cuh:142-154), then thrust::sort_by_key (stable by key).
void orcb_hash_and_index(const xdem_b_device_parameters* p, uint32_t* hash, uint32_t* index,
                         const double* pos4, uint32_t np) {
    uint32_t n = ids_below(np, (uint64_t)p->np + p->nsp);
    for (uint32_t i = 0; i < n; ++i) {
        int cx, cy, cz;
        cell_of(*p, ld(pos4 + 4 * i), cx, cy, cz);
        hash[i] = hash_of(*p, cx, cy, cz);
        index[i] = i;
    }
    std::vector<uint32_t> perm(np);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return hash[a] < hash[b]; });
    std::vector<uint32_t> h2(np), i2(np);
    for (uint32_t j = 0; j < np; ++j) {
        h2[j] = hash[perm[j]];
        i2[j] = index[perm[j]];
    }
    std::copy(h2.begin(), h2.end(), hash);
    std::copy(i2.begin(), i2.end(), index);
}

// cu_reorderDataAndFindCellStart — cu_dem_def.cu:71-90, kernel :156-196
void orcb_reorder(const xdem_b_device_parameters* p, const uint32_t* hash, const uint32_t* index,
                  uint32_t* cstart, uint32_t* cend, uint32_t* sorted, uint32_t np, uint32_t nc) {
    std::fill(cstart, cstart + nc, 0xFFFFFFFFu);
    const uint64_t tnp = (uint64_t)p->np + p->nsp;
    uint32_t n = ids_below(np, tnp);
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t h = hash[id];
        if (id == 0 || h != hash[id - 1]) {
            cstart[h] = id;
            if (id > 0) cend[hash[id - 1]] = id;
        }
        if ((uint64_t)id == tnp - 1) cend[h] = id + 1;
        sorted[id] = index[id];
    }
}

// calculate_p2p_kernel — cu_dem_impl.cuh:254-327
void orcb_p2p(const xdem_b_device_parameters* p, const double* pos4, const double* vel4,
              const double* acc4, const double* omega4, double* force3, double* moment3,
              const uint32_t* sorted, const uint32_t* cstart, const uint32_t* cend, uint32_t np) {
    uint32_t n = ids_below(np, p->np);
    for (uint32_t id = 0; id < n; ++id) {
        V3 ipos = ld(pos4 + 4 * id), ivel = ld(vel4 + 4 * id), iomega = ld(omega4 + 4 * id);
        double ir = pos4[4 * id + 3];
        int gx, gy, gz;
        cell_of(*p, ipos, gx, gy, gz);
        V3 F = acc4[4 * id + 3] * of(p->gravity), M = mk(0, 0, 0);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    uint32_t h = hash_of(*p, gx + x, gy + y, gz + z);
                    uint32_t s = cstart[h];
                    if (s == 0xFFFFFFFFu) continue;
                    uint32_t e = cend[h];
                    for (uint32_t j = s; j < e; ++j) {
                        uint32_t k = sorted[j];
                        if (id == k || k >= p->np) continue;
                        cal_force(*p, ir, pos4[4 * k + 3], ipos, ld(pos4 + 4 * k), ivel,
                                  ld(vel4 + 4 * k), iomega, ld(omega4 + 4 * k), F, M);
                    }
                }
        V3 f = ld(force3 + 3 * id), m = ld(moment3 + 3 * id);
        f += F;
        m += M;
        st(force3 + 3 * id, f);
        st(moment3 + 3 * id, m);
    }
}

// cube_hertzian_contact_force_kernel — cu_dem_impl.cuh:329-378
void orcb_cube(const xdem_b_device_parameters* p, const xdem_b_plane_info* planes, double kn,
               double vn, double ks, double vs, double mu, const double* pos4, const double* vel4,
               const double* omega4, double* force3, double* moment3, uint32_t np) {
    uint32_t n = ids_below(np, p->np);
    for (uint32_t id = 0; id < n; ++id) {
        double r = pos4[4 * id + 3];
        V3 ipos = ld(pos4 + 4 * id), ivel = ld(vel4 + 4 * id), iomega = ld(omega4 + 4 * id);
        V3 mf = mk(0, 0, 0), mm = mk(0, 0, 0);
        for (int i = 0; i < 6; i++) {
            const xdem_b_plane_info& pl = planes[i];
            V3 dp = ipos - of(pl.xw);
            V3 wp = mk(dot(dp, of(pl.u1)), dot(dp, of(pl.u2)), dot(dp, of(pl.uw)));
            if (std::fabs(wp.z) < r && (wp.x > 0 && wp.x < pl.l1) && (wp.y > 0 && wp.y < pl.l2)) {
                V3 unit = (double)(-sgn(dot(ipos - of(pl.xw), of(pl.uw)))) * (of(pl.uw) / length(of(pl.uw)));
                double cd = r - std::fabs(dot(ipos - of(pl.xw), unit));
                V3 dv = -(ivel + cross(iomega, r * unit));
                V3 sf = (-kn * std::pow(cd, 1.5) + vn * dot(dv, unit)) * unit;
                V3 e = dv - dot(dv, unit) * unit;
                double mag_e = length(e);
                if (mag_e) {
                    V3 s_hat = e / mag_e;
                    double ds = mag_e * p->dt;
                    V3 shear = std::fmin(ks * ds + vs * dot(dv, s_hat), mu * length(sf)) * s_hat;
                    mm += cross(r * unit, shear);
                }
                mf += sf;
            }
        }
        V3 f = ld(force3 + 3 * id), m = ld(moment3 + 3 * id);
        f += mf;
        m += mm;
        st(force3 + 3 * id, f);
        st(moment3 + 3 * id, m);
    }
}

// cu_shape_hertzian_contact_force — cu_dem_def.cu:140-232: clear F and M,
// face pass (:676-762), line/edge pass (:591-674), reduce6 of the forces.
void orcb_shape(const xdem_b_device_parameters* p, const xdem_b_mass_info* mi,
                const xdem_b_polygon* polys, const uint32_t* id_set, const uint32_t* poly_start,
                const uint32_t* poly_end, double kn, double vn, double ks, double vs, double mu,
                bool* is_line, const double* pos4, const double* vel4, const double* omega4,
                double* force3, double* moment3, uint32_t np, const uint32_t* sorted,
                const uint32_t* cstart, const uint32_t* cend, double out_bf[3]) {
    std::fill(force3, force3 + 3 * (size_t)np, 0.0);
    std::fill(moment3, moment3 + 3 * (size_t)np, 0.0);
    ShapeIn S{p, mi, polys, id_set, poly_start, poly_end, kn, vn, ks, vs, mu, pos4, vel4, omega4,
              sorted, cstart, cend};
    uint32_t n = ids_below(np, p->np);
    for (uint32_t id = 0; id < n; ++id) {
        is_line[id] = false;
        V3 F, M;
        if (shape_pass(S, id, true, F, M)) {
            V3 f = ld(force3 + 3 * id), m = ld(moment3 + 3 * id);
            f += F;
            m += M;
            st(force3 + 3 * id, f);
            st(moment3 + 3 * id, m);
            is_line[id] = true;
        }
    }
    for (uint32_t id = 0; id < n; ++id) {
        if (is_line[id]) continue;
        V3 F, M;
        if (shape_pass(S, id, false, F, M)) {
            is_line[id] = true;
            V3 f = ld(force3 + 3 * id), m = ld(moment3 + 3 * id);
            f += F;
            m += M;
            st(force3 + 3 * id, f);
            st(moment3 + 3 * id, m);
        }
    }
    reduce6_256(force3, np, out_bf);
}

// update_polygons_position — cu_dem_impl.cuh:764-780
void orcb_update_polygons_position(const double* mpos3, const double* A, const double* local3,
                                   uint32_t nv, double* pos3) {
    for (uint32_t id = 0; id < nv; ++id) {
        V3 lv = ld(local3 + 3 * id);
        double tx = A[0] * lv.x + A[1] * lv.y + A[2] * lv.z;
        double ty = A[3] * lv.x + A[4] * lv.y + A[5] * lv.z;
        double tz = A[6] * lv.x + A[7] * lv.y + A[8] * lv.z;
        pos3[3 * id] = mpos3[0] + tx;
        pos3[3 * id + 1] = mpos3[1] + ty;
        pos3[3 * id + 2] = mpos3[2] + tz;
    }
}

// update_polygons_information — cu_dem_impl.cuh:798-818
void orcb_update_polygons_information(xdem_b_polygon* polys, const uint32_t* indice3,
                                      const double* pos3, uint32_t ni) {
    for (uint32_t id = 0; id < ni; ++id) {
        const uint32_t* ix = indice3 + 3 * id;
        V3 P = ld(pos3 + 3 * ix[0]), Q = ld(pos3 + 3 * ix[1]), R = ld(pos3 + 3 * ix[2]);
        V3 Vv = Q - P, W = R - P, N = cross(Vv, W);
        polys[id].P = xdem_double3{P.x, P.y, P.z};
        polys[id].Q = xdem_double3{Q.x, Q.y, Q.z};
        polys[id].R = xdem_double3{R.x, R.y, R.z};
        polys[id].V = xdem_double3{Vv.x, Vv.y, Vv.z};
        polys[id].W = xdem_double3{W.x, W.y, W.z};
        polys[id].N = xdem_double3{N.x, N.y, N.z};
    }
}

}  // extern "C"
