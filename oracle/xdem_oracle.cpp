// xdem_oracle.cpp — CPU FP64 restatement of the reference DEM contact step.
//
// TEST INFRASTRUCTURE: the checker for libxdem.so and the timed CPU baseline.
// Nothing in the product links or calls this file.
//
// Build: g++ -O2 -ffp-contract=off -fopenmp -fPIC -shared (oracle/Makefile).
// -ffp-contract=off: every a*b+c below is two rounded operations, as the
// restated expressions are written; the HIP kernels are compiled the same way.
//
// Semantics.  ORC_GPU restates include/mphysics_cuda_impl.cuh (the live CUDA
// kernels); ORC_HOST restates the serial CPU path.  Where they differ the code
// below says which line each branch follows (the differences A1-A10 of the
// reference analysis: torque lever arm, plane accumulation, plane relative
// velocity, rolling switch, sign(0), abs(floor), cell-scan read, FMA, Verlet
// rounding).
//
// Deliberate departures from the reference (defects not reproduced):
//  * wall/triangle cohesion: the GPU kernels pass ri = rj = 0 to cohesionForce,
//    which yields 0/0 = NaN for coh != 0 (mphysics_cuda_impl.cuh:609,614 ->
//    :271-273).  Here a wall/triangle contact has no cohesion term.
//  * cylinder reaction: every thread writes *mf = 0 (a race with thread 0,
//    :732-733); here mf[id] = -Fn for every id.
//  * triangle reaction sums dpmi[].force/moment: racy and never zeroed in the
//    reference (:1010-1011); here zeroed per call and summed in particle order.
//  * reductionD3 with n < 512 reads uninitialised shared memory in the
//    reference; here missing lanes count as zero.
//  * degenerate triangle (all three circumcentre denominators <= 1e-13):
//    t is uninitialised in the reference (mphysics_cuda_impl.cuh:1938); here 0.
//  * the CPU cell scan reads cell_id[n] one past the end
//    (neighborhood_cell.cpp:73-76); here the exact run-length scan.
//
// Parity pins (tests/test_oracle_*.py): hash = closed form + exhaustive tiny
// grids + cell faces; sort = std::stable_sort definition and brute force; cell
// ranges = invariants; pair list = O(N^2) brute force; forces = two-sphere
// closed forms, Appendix-C constants, sphere-on-plane static overlap; closest
// point = brute-force barycentric minimisation; Verlet = exact parabola;
// reduction = exact integer-valued sums; circumcentre = equidistance; drum =
// static shell and cap overlaps, spinning-wall stick and Coulomb-saturated
// tangential closed forms, drive torque, invariance under a rotated axis;
// motion pose = rotation/translation closed forms, composition order and the
// active-step count.
//
// Rotating drum and prescribed motion have no reference kernel: orc_drum
// applies the cylinder kernel's contact evaluation (mphysics_cuda_impl.cuh:
// 720-774) to the inside of a spinning shell and its end caps; the motion
// follows the geometry motion condition (xDynamicsSolver.cpp:176-184,
// contact_particles_plane.cpp:33-49) and updatePolygonObjectData
// (mphysics_cuda_impl.cuh:1911-1956).
#include "xdem_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
    double x, y, z;
};
inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
inline V3 ld3(const double* p) { return V3{p[0], p[1], p[2]}; }
inline V3 ld4(const double* p) { return V3{p[0], p[1], p[2]}; }
inline V3 of(const xdem_double3& d) { return V3{d.x, d.y, d.z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline void operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
// dot: mphysics_cuda_impl.cuh:19-22 (left-to-right sum)
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
// cross: mphysics_cuda_impl.cuh:60-63
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline void st3(double* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// __umul24: product of the low 24 bits of each operand (low 32 bits kept).
inline uint32_t umul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

// getConstant — mphysics_cuda_impl.cuh:214-256; CPU twin contact.cpp:129-156
// (identical DHS formulas).
xdem_device_force_constant force_constant(int tcm, double ir, double jr, double im, double jm,
                                          const xdem_device_contact_property& cp) {
    xdem_device_force_constant dfc = {0, 0, 0, 0, 0, 0};
    double Meq = jm ? (im * jm) / (im + jm) : im;
    double Req = jr ? (ir * jr) / (ir + jr) : ir;
    double Eeq = (cp.Ei * cp.Ej) / (cp.Ei * (1 - cp.prj * cp.prj) + cp.Ej * (1 - cp.pri * cp.pri));
    if (tcm == 0) {
        double Geq = (cp.Gi * cp.Gj) / (cp.Gi * (2 - cp.prj) + cp.Gj * (2 - cp.pri));
        double ln_e = std::log(cp.rest);
        double xi = ln_e / std::sqrt(ln_e * ln_e + M_PI * M_PI);
        dfc.kn = (4.0 / 3.0) * Eeq * std::sqrt(Req);
        dfc.vn = -2.0 * std::sqrt(5.0 / 6.0) * xi * std::sqrt(dfc.kn * Meq);
        dfc.ks = 8.0 * Geq * std::sqrt(Req);
        dfc.vs = -2.0 * std::sqrt(5.0 / 6.0) * xi * std::sqrt(dfc.ks * Meq);
        dfc.mu = cp.fric;
        dfc.ms = cp.rfric;
    } else {
        double beta = (M_PI / std::log(cp.rest));
        dfc.kn = (4.0 / 3.0) * Eeq * std::sqrt(Req);
        dfc.vn = std::sqrt((4.0 * Meq * dfc.kn) / (1.0 + beta * beta));
        dfc.ks = dfc.kn * cp.sratio;
        dfc.vs = dfc.vn * cp.sratio;
        dfc.mu = cp.fric;
        dfc.ms = cp.rfric;
    }
    return dfc;
}

// Optional per-particle sum of contact-force magnitudes |Fn + Ft| (set by
// orc_set_force_magnitudes): the cancellation-safe scale of a per-particle
// force comparison.  Each particle's entry is written only by the loop
// iteration of that particle.
//
// A second sum (g_tmag) adds the magnitudes of the terms the contact law
// combines: |fsn| + |fca| + (|vn| + |vs|) V + |Ft|, V the summed magnitudes of
// the velocities entering the relative velocity (|v_i| + |w_i| r_i, plus the
// partner's |v_j| + |w_j| r_j or the wall point's speed).  Rounding in a
// lower-precision evaluation of the law is bounded relative to that sum; the
// net force |Fn + Ft| can be much smaller when the terms cancel.
double* g_fmag = nullptr;
double* g_tmag = nullptr;
inline void add_mag(uint32_t id, const V3& f, double terms = 0.0) {
    if (g_fmag) g_fmag[id] += length(f);
    if (g_tmag) g_tmag[id] += terms;
}
inline double term_scale(const xdem_device_force_constant& c, double cdist, double fca, double V,
                         const V3& Ft) {
    return std::fabs(c.kn * std::pow(cdist, 1.5)) + std::fabs(fca) +
           (std::fabs(c.vn) + std::fabs(c.vs)) * V + length(Ft);
}

// cohesionForce — mphysics_cuda_impl.cuh:259-279 (particle-particle only; walls
// and triangles get no cohesion term, see header).
double cohesion_force(double ri, double rj, double Ei, double Ej, double pri, double prj,
                      double coh, double Fn) {
    double cf = 0.0;
    if (coh) {
        double req = (ri * rj / (ri + rj));
        double Eeq = ((1.0 - pri * pri) / Ei) + ((1.0 - prj * prj) / Ej);
        double rcp = (3.0 * req * (-Fn)) / (4.0 * (1.0 / Eeq));
        double rc = std::pow(rcp, 1.0 / 3.0);
        double Ac = M_PI * rc * rc;
        cf = coh * Ac;
    }
    return cf;
}

// DHSModel — mphysics_cuda_impl.cuh:379-401.  Ft and M are in/out: when the
// tangential velocity is exactly zero they keep the caller's previous values
// (the kernels declare them outside the neighbour loop, :426-428, :945-949).
void dhs_model(const xdem_device_force_constant& c, double dt, double ir, double fca, double cdist,
               V3 dv, V3 unit, V3& Ft, V3& Fn, V3& M) {
    double fsn = -c.kn * std::pow(cdist, 1.5);
    double fdn = c.vn * dot(dv, unit);
    Fn = (fsn + fca + fdn) * unit;
    V3 e = dv - dot(dv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        V3 sh = e / mag_e;
        double ds = mag_e * dt;
        Ft = std::fmin(c.ks * ds + c.vs * (dot(dv, sh)), c.mu * length(Fn)) * sh;
        M = cross(ir * unit, Ft);
    }
}

// HMCModel — mphysics_cuda_impl.cuh:347-377.
void hmc_model(const xdem_device_force_constant& c, double dt, double ir, double fca, double rcon,
               double cdist, V3 iomega, V3 dv, V3 unit, V3& Ft, V3& Fn, V3& M) {
    double fsn = -c.kn * std::pow(cdist, 1.5);
    double fdn = c.vn * dot(dv, unit);
    Fn = (fsn + fca + fdn) * unit;
    V3 e = dv - dot(dv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        V3 s_hat = -(e / mag_e);
        double ds = mag_e * dt;
        double fst = -c.ks * ds;
        double fdt = c.vs * dot(dv, s_hat);
        Ft = (fst + fdt) * s_hat;
        if (length(Ft) >= c.mu * length(Fn)) Ft = c.mu * fsn * s_hat;
        M = cross(ir * unit, Ft);
        if (length(iomega)) {
            V3 on = iomega / length(iomega);
            M += c.ms * fsn * rcon * on;
        }
    }
}

// Force-law dispatch for the GPU semantics (TCM switch of the kernels).
inline void gpu_law(int tcm, const xdem_device_force_constant& c, double dt, double ir, double fca,
                    double rcon, double cdist, V3 iomega, V3 dv, V3 unit, V3& Ft, V3& Fn, V3& M) {
    if (tcm == 0)
        hmc_model(c, dt, ir, fca, rcon, cdist, iomega, dv, unit, Ft, Fn, M);
    else
        dhs_model(c, dt, ir, fca, cdist, dv, unit, Ft, Fn, M);
}

// contact::DHSModel — source/collision/contact.cpp:20-46 (CPU path): F is
// assigned Fn + Ft (Ft zero when |e| == 0), M = cp x Ft only when |e| != 0,
// with cp = rcon * unit (the A1 lever arm).
void host_dhs(const xdem_device_force_constant& c, double dt, double fca, double cdist, V3 cp,
              V3 dv, V3 unit, V3& F, V3& M) {
    V3 Fn, Ft = mk(0, 0, 0);
    double fsn = (-c.kn * std::pow(cdist, 1.5));
    double fsd = c.vn * dot(dv, unit);
    Fn = (fsn + fca + fsd) * unit;
    V3 e = dv - dot(dv, unit) * unit;
    double mag_e = length(e);
    if (mag_e) {
        V3 s_hat = e / mag_e;
        double ds = mag_e * dt;
        double ft1 = c.ks * ds + c.vs * dot(dv, s_hat);
        double ft2 = c.mu * length(Fn);
        Ft = std::fmin(ft1, ft2) * s_hat;
        M = cross(cp, Ft);
    }
    F = Fn + Ft;
}

// contact::cohesionForce — contact.cpp:6-18 with coh_r = R*, coh_e = sum (1-nu^2)/E
double host_cohesion(double coh, double coh_r, double coh_e, double Fn) {
    double cf = 0.0;
    if (coh) {
        double rcp = (3.0 * coh_r * (-Fn)) / (4.0 * (1.0 / coh_e));
        double rc = std::pow(rcp, 1.0 / 3.0);
        double Ac = M_PI * rc * rc;
        cf = coh * Ac;
    }
    return cf;
}

inline int gpu_sign(double L) { return L < 0 ? -1 : 1; }  // mphysics_cuda_impl.cuh:14-17
// algebraMath.h:6 `sign` macro as expanded in `-sign(x)` (contact_particles_plane.cpp:137):
// -x <= 0 ? (x == 0 ? 0 : -1) : 1, i.e. the negated sign with sign(0) = 0.
inline int host_neg_sign(double x) { return -x <= 0 ? (x == 0 ? 0 : -1) : 1; }

void sort_cells(uint32_t* hash, uint32_t* index, uint32_t n, uint32_t ncell, uint32_t* cstart,
                uint32_t* cend, uint32_t* sorted) {
    // thrust::sort_by_key (stable LSD radix) == stable sort by key.
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return hash[a] < hash[b]; });
    std::vector<uint32_t> h2(n), i2(n);
    for (uint32_t j = 0; j < n; ++j) {
        h2[j] = hash[perm[j]];
        i2[j] = index[perm[j]];
    }
    std::memcpy(hash, h2.data(), sizeof(uint32_t) * n);
    std::memcpy(index, i2.data(), sizeof(uint32_t) * n);
    // cudaMemset(cstart, 0xff), cudaMemset(cend, 0): mphysics_cuda_def.cu:97-98
    std::fill(cstart, cstart + ncell, 0xFFFFFFFFu);
    std::fill(cend, cend + ncell, 0u);
    // reorderDataAndFindCellStart_kernel — mphysics_cuda_impl.cuh:166-212
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t h = hash[id];
        if (id == 0 || h != hash[id - 1]) {
            cstart[h] = id;
            if (id > 0) cend[hash[id - 1]] = id;
        }
        if (id == n - 1) cend[h] = id + 1;
        sorted[id] = index[id];
    }
}

}  // namespace

extern "C" {

// calcGridPos (mphysics_cuda_impl.cuh:75-83) / grid_base::getCellNumber
// (grid_base.cpp:139-146, which takes abs() of the floored value: A6).
void orc_cell_of(const xdem_device_parameters* p, int semantics, double x, double y, double z,
                 int32_t out[3]) {
    double fx = std::floor((x - p->world_origin.x) / p->cell_size);
    double fy = std::floor((y - p->world_origin.y) / p->cell_size);
    double fz = std::floor((z - p->world_origin.z) / p->cell_size);
    if (semantics == ORC_HOST) {
        fx = std::fabs(fx);
        fy = std::fabs(fy);
        fz = std::fabs(fz);
    }
    out[0] = (int32_t)fx;
    out[1] = (int32_t)fy;
    out[2] = (int32_t)fz;
}

// calcGridHash (mphysics_cuda_impl.cuh:65-72) / grid_base::getHash (grid_base.cpp:148-155)
uint32_t orc_hash_of(const xdem_device_parameters* p, const int32_t c[3]) {
    uint32_t gx = p->grid_size.x, gy = p->grid_size.y, gz = p->grid_size.z;
    uint32_t x = (uint32_t)c[0] & (gx - 1);
    uint32_t y = (uint32_t)c[1] & (gy - 1);
    uint32_t z = (uint32_t)c[2] & (gz - 1);
    return umul24(umul24(z, gy), gx) + umul24(y, gx) + x;
}

// calculateHashAndIndex_kernel (:137-151) and ...ForPolygonSphere_kernel (:153-164)
void orc_hash_and_index(const xdem_device_parameters* p, int semantics, const double* pos4,
                        uint32_t n, uint32_t off, uint32_t* hash, uint32_t* index) {
    for (uint32_t i = 0; i < n; ++i) {
        int32_t c[3];
        orc_cell_of(p, semantics, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], c);
        hash[off + i] = orc_hash_of(p, c);
        index[off + i] = off + i;
    }
}

void orc_reorder_and_find_cell_start(uint32_t* hash, uint32_t* index, uint32_t n, uint32_t ncell,
                                     uint32_t* cstart, uint32_t* cend, uint32_t* sorted) {
    sort_cells(hash, index, n, ncell, cstart, cend, sorted);
}

double orc_cohesion_force(double ri, double rj, double Ei, double Ej, double pri, double prj,
                          double coh, double Fn) {
    return cohesion_force(ri, rj, Ei, Ej, pri, prj, coh, Fn);
}

double orc_host_cohesion(double coh, double coh_r, double coh_e, double Fn) {
    return host_cohesion(coh, coh_r, coh_e, Fn);
}

void orc_set_force_magnitudes(double* fmag, double* tmag) {
    g_fmag = fmag;
    g_tmag = tmag;
}

void orc_force_constant(int32_t tcm, double ir, double jr, double im, double jm,
                        const xdem_device_contact_property* cp, xdem_device_force_constant* out) {
    *out = force_constant(tcm, ir, jr, im, jm, *cp);
}

}  // extern "C"

namespace {

// One particle of calculate_p2p_kernel<TCM> (mphysics_cuda_impl.cuh:403-485),
// or of the CPU 27-cell loop (contactManager.cpp:235-272 ->
// contact_particles_particles.cpp:53-89) in HOST semantics.  Returns the
// contribution to add (sumF, sumM); appends pairs when requested.
void p2p_one(const xdem_device_parameters* p, int tcm, int sem, uint32_t id, const double* pos4,
             const double* vel, const double* omega, const double* mass, const uint32_t* sorted,
             const uint32_t* cstart, const uint32_t* cend, const xdem_device_contact_property& cp,
             uint32_t np, V3& sumF, V3& sumM, uint32_t* pairs, uint64_t max_pairs,
             uint64_t* npairs) {
    V3 ipos = ld4(pos4 + 4 * id);
    double ir = pos4[4 * id + 3];
    V3 ivel = ld3(vel + 3 * id), iomega = ld3(omega + 3 * id);
    double im = mass[id];
    int32_t g[3];
    orc_cell_of(p, sem, ipos.x, ipos.y, ipos.z, g);
    V3 Ft = mk(0, 0, 0), Fn = mk(0, 0, 0), M = mk(0, 0, 0);
    sumF = mk(0, 0, 0);
    sumM = mk(0, 0, 0);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                int32_t nb[3] = {g[0] + x, g[1] + y, g[2] + z};
                uint32_t h = orc_hash_of(p, nb);
                uint32_t s = cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = sorted[j];
                    if (id == k || k >= np) continue;
                    V3 jpos = ld4(pos4 + 4 * k);
                    double jr = pos4[4 * k + 3];
                    V3 jvel = ld3(vel + 3 * k), jomega = ld3(omega + 3 * k);
                    double jm = mass[k];
                    V3 rp = mk(jpos.x - ipos.x, jpos.y - ipos.y, jpos.z - ipos.z);
                    double dist = length(rp);
                    double cdist = (ir + jr) - dist;
                    if (!(cdist > 0)) continue;
                    if (pairs && *npairs < max_pairs) {
                        pairs[2 * *npairs] = id;
                        pairs[2 * *npairs + 1] = k;
                    }
                    ++*npairs;
                    double rcon = ir - 0.5 * cdist;
                    V3 unit = rp / dist;
                    xdem_device_force_constant c = force_constant(tcm, ir, jr, im, jm, cp);
                    if (sem == ORC_HOST) {
                        // contact_particles_particles.cpp:68-88 (rolling terms always: A4;
                        // lever arm rcon: A1)
                        V3 cpt = rcon * unit;
                        V3 rv = jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit));
                        double coh_e = ((1.0 - cp.pri * cp.pri) / cp.Ei) + ((1.0 - cp.prj * cp.prj) / cp.Ej);
                        double Req = jr ? (ir * jr) / (ir + jr) : ir;
                        double fsn = (-c.kn * std::pow(cdist, 1.5));
                        double fca = host_cohesion(cp.coh, Req, coh_e, fsn);
                        V3 mf = mk(0, 0, 0), mm = mk(0, 0, 0);
                        host_dhs(c, p->dt, fca, cdist, cpt, rv, unit, mf, mm);
                        add_mag(id, mf,
                                term_scale(c, cdist, fca,
                                           length(ivel) + length(iomega) * ir + length(jvel) +
                                               length(jomega) * jr,
                                           mf));
                        sumF += mf;
                        sumM += mm;
                    } else {
                        V3 rv = p->rollingCondition
                                    ? jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit))
                                    : (jvel - ivel);
                        double fsn = -c.kn * std::pow(cdist, 1.5);
                        double fca = cohesion_force(ir, jr, cp.Ei, cp.Ej, cp.pri, cp.prj, cp.coh, fsn);
                        gpu_law(tcm, c, p->dt, ir, fca, rcon, cdist, iomega, rv, unit, Ft, Fn, M);
                        add_mag(id, Fn + Ft,
                                term_scale(c, cdist, fca,
                                           length(ivel) + length(iomega) * ir + length(jvel) +
                                               length(jomega) * jr,
                                           Ft));
                        sumF += Fn + Ft;
                        sumM += M;
                    }
                }
            }
}

}  // namespace

extern "C" {

uint64_t orc_p2p(const xdem_device_parameters* p, int32_t tcm, int semantics, const double* pos4,
                 const double* vel, const double* omega, double* force, double* moment,
                 const double* mass, const uint32_t* sorted, const uint32_t* cstart,
                 const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np,
                 uint32_t* pairs, uint64_t max_pairs) {
    uint64_t npairs = 0;
    for (uint32_t id = 0; id < np; ++id) {
        V3 sF, sM;
        p2p_one(p, tcm, semantics, id, pos4, vel, omega, mass, sorted, cstart, cend, *cp, np, sF,
                sM, pairs, max_pairs, &npairs);
        double* f = force + 3 * id;
        double* m = moment + 3 * id;
        f[0] += sF.x; f[1] += sF.y; f[2] += sF.z;
        m[0] += sM.x; m[1] += sM.y; m[2] += sM.z;
    }
    return npairs;
}

uint64_t orc_pairs_bruteforce(const double* pos4, uint32_t np, uint32_t* pairs,
                              uint64_t max_pairs) {
    uint64_t n = 0;
    for (uint32_t i = 0; i < np; ++i)
        for (uint32_t k = 0; k < np; ++k) {
            if (i == k) continue;
            V3 rp = mk(pos4[4 * k] - pos4[4 * i], pos4[4 * k + 1] - pos4[4 * i + 1],
                       pos4[4 * k + 2] - pos4[4 * i + 2]);
            double cdist = (pos4[4 * i + 3] + pos4[4 * k + 3]) - length(rp);
            if (cdist > 0) {
                if (pairs && n < max_pairs) {
                    pairs[2 * n] = i;
                    pairs[2 * n + 1] = k;
                }
                ++n;
            }
        }
    return n;
}

// particle_plane_contact_detection — mphysics_cuda_impl.cuh:487-570 (GPU) /
// contact_particles_plane.cpp:122-207 (HOST: sign(0) = 0, A5).  Branch order,
// including the unreachable seventh branch (:549), is kept.
double orc_plane_detect(const xdem_device_plane_info* pe, int semantics, const double xp_[3],
                        const double wp_[3], double u_[3], double r) {
    V3 xp = ld3(xp_), wp = ld3(wp_), u = ld3(u_);
    V3 xw = of(pe->xw), w2 = of(pe->w2), w3 = of(pe->w3), w4 = of(pe->w4), uw = of(pe->uw);
    double a_l1 = std::pow(wp.x - pe->l1, 2.0);
    double b_l2 = std::pow(wp.y - pe->l2, 2.0);
    double sqa = wp.x * wp.x;
    double sqb = wp.y * wp.y;
    double sqc = wp.z * wp.z;
    double sqr = r * r;
    double res = -1.0;
    if (std::fabs(wp.z) < r && (wp.x > 0 && wp.x < pe->l1) && (wp.y > 0 && wp.y < pe->l2)) {
        V3 dp = xp - xw;
        V3 uu = uw / length(uw);
        int pp = semantics == ORC_HOST ? host_neg_sign(dot(dp, uw)) : -gpu_sign(dot(dp, uw));
        u = (double)pp * uu;
        res = r - std::fabs(dot(dp, u));
        st3(u_, u);
        return res;
    }
    if (wp.x < 0 && wp.y < 0 && (sqa + sqb + sqc) < sqr) {
        V3 Xsw = xp - xw;
        double h = length(Xsw);
        u = Xsw / h;
        res = r - h;
    } else if (wp.x > pe->l1 && wp.y < 0 && (a_l1 + sqb + sqc) < sqr) {
        V3 Xsw = xp - w2;
        double h = length(Xsw);
        u = Xsw / h;
        res = r - h;
    } else if (wp.x > pe->l1 && wp.y > pe->l2 && (a_l1 + b_l2 + sqc) < sqr) {
        V3 Xsw = xp - w3;
        double h = length(Xsw);
        u = Xsw / h;
        res = r - h;
    } else if (wp.x < 0 && wp.y > pe->l2 && (sqa + b_l2 + sqc) < sqr) {
        V3 Xsw = xp - w4;
        double h = length(Xsw);
        u = Xsw / h;
        res = r - h;
    } else if ((wp.x > 0 && wp.x < pe->l1) && wp.y < 0 && (sqb + sqc) < sqr) {
        V3 Xsw = xp - xw;
        V3 wj_wi = w2 - xw;
        V3 us = wj_wi / length(wj_wi);
        V3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        res = r - h;
    } else if ((wp.x > 0 && wp.x < pe->l1) && wp.y > pe->l2 && (b_l2 + sqc) < sqr) {
        V3 Xsw = xp - w4;
        V3 wj_wi = w3 - w4;
        V3 us = wj_wi / length(wj_wi);
        V3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        res = r - h;
    } else if ((wp.x > 0 && wp.y < pe->l2) && wp.x < 0 && (sqr + sqc) < sqr) {
        V3 Xsw = xp - xw;
        V3 wj_wi = w4 - xw;
        V3 us = wj_wi / length(wj_wi);
        V3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        res = r - h;
    } else if ((wp.x > 0 && wp.y < pe->l2) && wp.x > pe->l1 && (a_l1 + sqc) < sqr) {
        V3 Xsw = xp - w2;
        V3 wj_wi = w3 - w2;
        V3 us = wj_wi / length(wj_wi);
        V3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        res = r - h;
    }
    st3(u_, u);
    return res;
}

// plane_contact_force_kernel<TCM> — mphysics_cuda_impl.cuh:572-621.  The force
// law is called with ir = 0 (:609,614), so plane contacts add no torque.
void orc_plane(const xdem_device_parameters* p, int32_t tcm, const xdem_device_plane_info* pl,
               const double* pos4, const double* vel, const double* omega, double* force,
               double* moment, const double* mass, uint32_t np,
               const xdem_device_contact_property* cp) {
    for (uint32_t id = 0; id < np; ++id) {
        double m = mass[id];
        V3 ipos3 = ld4(pos4 + 4 * id);
        double r = pos4[4 * id + 3];
        V3 ivel = ld3(vel + 3 * id), iomega = ld3(omega + 3 * id);
        V3 Fn = mk(0, 0, 0), Ft = mk(0, 0, 0), M = mk(0, 0, 0);
        V3 dp = ipos3 - of(pl->xw);
        double unit[3] = {0, 0, 0};
        double wp[3] = {dot(dp, of(pl->u1)), dot(dp, of(pl->u2)), dot(dp, of(pl->uw))};
        double xp[3] = {ipos3.x, ipos3.y, ipos3.z};
        double cdist = orc_plane_detect(pl, ORC_GPU, xp, wp, unit, r);
        if (cdist > 0) {
            double rcon = r - 0.5 * cdist;
            V3 un = ld3(unit);
            V3 dv = p->rollingCondition ? -(ivel + cross(iomega, r * un)) : -ivel;
            xdem_device_force_constant c = force_constant(tcm, r, 0.0, m, 0.0, *cp);
            gpu_law(tcm, c, p->dt, 0.0, 0.0, rcon, cdist, iomega, dv, un, Ft, Fn, M);
            add_mag(id, Fn + Ft, term_scale(c, cdist, 0.0, length(ivel) + length(iomega) * r, Ft));
        }
        V3 s = Fn + Ft;
        double* f = force + 3 * id;
        double* mo = moment + 3 * id;
        f[0] += s.x; f[1] += s.y; f[2] += s.z;
        mo[0] += M.x; mo[1] += M.y; mo[2] += M.z;
    }
}

}  // extern "C"

namespace {
// makeTFM_1..3, toLocal, toGlobal — mphysics_cuda_impl.cuh:624-650
V3 tfm1(const xdem_double4& ep) {
    return mk(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w),
              2.0 * (ep.y * ep.w + ep.x * ep.z));
}
V3 tfm2(const xdem_double4& ep) {
    return mk(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5),
              2.0 * (ep.z * ep.w - ep.x * ep.y));
}
V3 tfm3(const xdem_double4& ep) {
    return mk(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y),
              2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
}
V3 to_local(V3 A1, V3 A2, V3 A3, V3 v) {
    return mk(A1.x * v.x + A2.x * v.y + A3.x * v.z, A1.y * v.x + A2.y * v.y + A3.y * v.z,
              A1.z * v.x + A2.z * v.y + A3.z * v.z);
}
V3 to_global(V3 A1, V3 A2, V3 A3, V3 v) { return mk(dot(A1, v), dot(A2, v), dot(A3, v)); }
}  // namespace

extern "C" {

// particle_cylinder_contact_detection — mphysics_cuda_impl.cuh:652-718.  The
// reference returns float, so the penetration is rounded to FP32 and widened.
double orc_cylinder_detect(const xdem_device_cylinder_info* cy, const double pt[4], double u_[3],
                           double cp_[3]) {
    V3 u = ld3(u_), cp = ld3(cp_);
    double dist = -1.0;
    V3 pbase = of(cy->pbase), ptop = of(cy->ptop), origin = of(cy->origin);
    V3 ab = mk(ptop.x - pbase.x, ptop.y - pbase.y, ptop.z - pbase.z);
    V3 p = mk(pt[0], pt[1], pt[2]);
    double w = pt[3];
    double t = dot(p - pbase, ab) / dot(ab, ab);
    V3 _cp = mk(0.0, 0.0, 0.0);
    float result = -1.0f;
    if (t >= 0 && t <= 1) {
        _cp = pbase + t * ab;
        dist = length(p - _cp);
        u = (_cp - p) / dist;
        cp = _cp - cy->rbase * u;
        result = (float)(cy->rtop + w - dist);
    } else {
        _cp = pbase + t * ab;
        dist = length(p - _cp);
        if (dist < cy->rbase) {
            V3 OtoCp = origin - _cp;
            double OtoCp_ = length(OtoCp);
            u = OtoCp / OtoCp_;
            cp = _cp - cy->rbase * u;
            result = (float)(cy->len * 0.5 + w - OtoCp_);
        } else {
            V3 A_1 = tfm1(cy->ep), A_2 = tfm2(cy->ep), A_3 = tfm3(cy->ep);
            V3 _at = p - ptop;
            V3 at = to_local(A_1, A_2, A_3, _at);
            cp = ptop;
            if (std::fabs(at.y) > cy->len) {
                _at = p - pbase;
                at = to_local(A_1, A_2, A_3, _at);
                cp = pbase;
            }
            double pi = std::atan(at.x / at.z);
            if (pi < 0 && at.z < 0)
                _cp.x = cy->rbase * std::sin(-pi);
            else if (pi > 0 && at.x < 0 && at.z < 0)
                _cp.x = cy->rbase * std::sin(-pi);
            else
                _cp.x = cy->rbase * std::sin(pi);
            _cp.z = cy->rbase * std::cos(pi);
            if (at.z < 0 && _cp.z > 0)
                _cp.z = -_cp.z;
            else if (at.z > 0 && _cp.z < 0)
                _cp.z = -_cp.z;
            _cp.y = 0.;
            cp = cp + to_global(A_1, A_2, A_3, _cp);
            V3 disVec = cp - p;
            dist = length(disVec);
            u = disVec / dist;
            if (dist < w) result = (float)(w - dist);
        }
    }
    st3(u_, u);
    st3(cp_, cp);
    return (double)result;
}

// cylinder_hertzian_contact_force_kernel<TCM> — mphysics_cuda_impl.cuh:720-774.
void orc_cylinder(const xdem_device_parameters* p, int32_t tcm, const xdem_device_cylinder_info* cy,
                  const double* pos4, const double* vel, const double* omega, double* force,
                  double* moment, const double* mass, uint32_t np,
                  const xdem_device_contact_property* cp, const double* mpos3, double* mf3,
                  double* mm3) {
    V3 mp = mpos3 ? ld3(mpos3) : mk(0, 0, 0);
    for (uint32_t id = 0; id < np; ++id) {
        double im = mass[id];
        const double* ip = pos4 + 4 * id;
        V3 ivel = ld3(vel + 3 * id), iomega = ld3(omega + 3 * id);
        double unit[3] = {0, 0, 0}, cpt[3] = {0, 0, 0};
        double cdist = orc_cylinder_detect(cy, ip, unit, cpt);
        V3 si = ld3(cpt) - mp;
        V3 cy2cp = ld3(cpt) - of(cy->origin);
        V3 Ft = mk(0, 0, 0), Fn = mk(0, 0, 0), M = mk(0, 0, 0);
        if (cdist > 0) {
            double rcon = ip[3] - 0.5 * cdist;
            V3 un = ld3(unit);
            V3 dv = of(cy->vel) + cross(of(cy->omega), cy2cp) - (ivel + cross(iomega, ip[3] * un));
            xdem_device_force_constant c = force_constant(tcm, ip[3], 0, im, 0, *cp);
            gpu_law(tcm, c, p->dt, 0.0, 0.0, rcon, cdist, iomega, dv, un, Ft, Fn, M);
            add_mag(id, Fn + Ft,
                    term_scale(c, cdist, 0.0,
                               length(ivel) + length(iomega) * ip[3] +
                                   length(of(cy->vel) + cross(of(cy->omega), cy2cp)),
                               Ft));
        }
        V3 sum_f = Fn + Ft;
        double* f = force + 3 * id;
        double* mo = moment + 3 * id;
        f[0] += sum_f.x; f[1] += sum_f.y; f[2] += sum_f.z;
        mo[0] += M.x; mo[1] += M.y; mo[2] += M.z;
        if (mf3) st3(mf3 + 3 * id, -(Fn));
        if (mm3) st3(mm3 + 3 * id, cross(si, -Fn));
    }
}

namespace {
// One drum contact: the cylinder kernel's evaluation (ir = 0, wall point
// velocity minus the particle's contact-point velocity) on an inner surface.
void drum_one(const xdem_device_parameters* p, int32_t tcm, const xdem_drum_info* dr, uint32_t id,
              double r, double im, V3 iomega, V3 ivel, V3 unit, V3 cpt, double cdist,
              const xdem_device_contact_property& cp, V3& F, V3& M, V3& R, V3& RM) {
    V3 o = of(dr->origin);
    V3 wv = dr->omega * cross(of(dr->axis), cpt - o);
    V3 dv = wv - (ivel + cross(iomega, r * unit));
    double rcon = r - 0.5 * cdist;
    V3 Ft = mk(0, 0, 0), Fn = mk(0, 0, 0), Mc = mk(0, 0, 0);
    xdem_device_force_constant c = force_constant(tcm, r, 0, im, 0, cp);
    gpu_law(tcm, c, p->dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, Mc);
    V3 f = Fn + Ft;
    add_mag(id, f, term_scale(c, cdist, 0.0, length(ivel) + length(iomega) * r + length(wv), Ft));
    F += f;
    M += Mc;
    R += -f;
    RM += cross(cpt - o, -f);
}
}  // namespace

// Rotating drum: shell contact (penetration r + rho - R, rho = distance from
// the axis) then cap contact (r - (half_length - |a|)), for centres between
// the caps; their sum is added once.
void orc_drum(const xdem_device_parameters* p, int32_t tcm, const xdem_drum_info* dr,
              const double* pos4, const double* vel, const double* omega, double* force,
              double* moment, const double* mass, uint32_t np,
              const xdem_device_contact_property* cp, double* mf3, double* mm3) {
    V3 ax = of(dr->axis);
    for (uint32_t id = 0; id < np; ++id) {
        const double* ip = pos4 + 4 * (size_t)id;
        double r = ip[3];
        V3 d = ld4(ip) - of(dr->origin);
        double a = dot(d, ax);
        V3 q = d - a * ax;
        double rho = std::sqrt(dot(q, q));
        V3 F = mk(0, 0, 0), M = mk(0, 0, 0), R = mk(0, 0, 0), RM = mk(0, 0, 0);
        if (std::fabs(a) <= dr->half_length) {
            double im = mass[id];
            V3 ivel = ld3(vel + 3 * (size_t)id), iomega = ld3(omega + 3 * (size_t)id);
            double cs = r + rho - dr->radius;
            if (cs > 0 && rho > 0) {
                V3 u = q / rho;
                V3 cpt = of(dr->origin) + a * ax + dr->radius * u;
                drum_one(p, tcm, dr, id, r, im, iomega, ivel, u, cpt, cs, *cp, F, M, R, RM);
            }
            double cc = r - (dr->half_length - std::fabs(a));
            if (dr->caps && cc > 0 && rho < dr->radius) {
                double sg = a >= 0 ? 1.0 : -1.0;
                V3 u = sg * ax;
                V3 cpt = of(dr->origin) + (sg * dr->half_length) * ax + q;
                drum_one(p, tcm, dr, id, r, im, iomega, ivel, u, cpt, cc, *cp, F, M, R, RM);
            }
        }
        double* f = force + 3 * (size_t)id;
        double* mo = moment + 3 * (size_t)id;
        f[0] += F.x; f[1] += F.y; f[2] += F.z;
        mo[0] += M.x; mo[1] += M.y; mo[2] += M.z;
        if (mf3) st3(mf3 + 3 * (size_t)id, R);
        if (mm3) st3(mm3 + 3 * (size_t)id, RM);
    }
}

// closestPtPointTriangle — mphysics_cuda_impl.cuh:837-899 (Ericson, Real-Time
// Collision Detection, 5.1.5).
void orc_closest_pt_triangle(const xdem_device_polygon_info* t, const double p_[3], double out[3]) {
    V3 a = of(t->P), b = of(t->Q), c = of(t->R), p = ld3(p_);
    V3 ab = b - a, ac = c - a, ap = p - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { st3(out, a); return; }
    V3 bp = p - b;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { st3(out, b); return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        st3(out, a + v * ab);
        return;
    }
    V3 cp = p - c;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { st3(out, c); return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        st3(out, a + w * ac);
        return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        st3(out, b + w * (c - b));
        return;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom;
    double w = vc * denom;
    st3(out, a + v * ab + w * ac);
}

}  // extern "C"

namespace {
// One particle of particle_polygonObject_collision_kernel<1>
// (mphysics_cuda_impl.cuh:923-1021).  Reaction contributions go to rf/rm
// (per object) when non-null.
void poly_one(const xdem_device_parameters* p, int tcm, uint32_t id,
              const xdem_device_polygon_info* dpi, const xdem_device_polygon_mass_info* dpmi,
              const double* pos4, const double* vel, const double* omega, const double* mass,
              const uint32_t* sorted, const uint32_t* cstart, const uint32_t* cend,
              const xdem_device_contact_property* cp, uint32_t np, V3& sum_force,
              V3& sum_moment, V3* rf, V3* rm) {
    double im = mass[id];
    V3 ipos = ld4(pos4 + 4 * id);
    V3 ivel = ld3(vel + 3 * id), iomega = ld3(omega + 3 * id);
    V3 unit = mk(0, 0, 0);
    int32_t g[3];
    orc_cell_of(p, ORC_GPU, ipos.x, ipos.y, ipos.z, g);
    double ir = pos4[4 * id + 3];
    V3 M = mk(0, 0, 0), Fn = mk(0, 0, 0), Ft = mk(0, 0, 0);
    sum_force = mk(0, 0, 0);
    sum_moment = mk(0, 0, 0);
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                int32_t nb[3] = {g[0] + x, g[1] + y, g[2] + z};
                uint32_t h = orc_hash_of(p, nb);
                uint32_t s = cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = sorted[j];
                    if (k < np) continue;
                    k -= np;
                    uint32_t pidx = (uint32_t)dpi[k].id;
                    const xdem_device_contact_property& cmp = cp[pidx];
                    const xdem_device_polygon_mass_info& pmi = dpmi[pidx];
                    double ipa[3] = {ipos.x, ipos.y, ipos.z}, cpa[3];
                    orc_closest_pt_triangle(&dpi[k], ipa, cpa);
                    V3 cpt = ld3(cpa);
                    V3 po2cp = cpt - of(pmi.origin);
                    V3 distVec = ipos - cpt;
                    double dist = length(distVec);
                    double cdist = ir - dist;
                    Fn = mk(0.0, 0.0, 0.0);
                    if (cdist > 0) {
                        V3 qp = of(dpi[k].Q) - of(dpi[k].P);
                        V3 rp = of(dpi[k].R) - of(dpi[k].P);
                        double rcon = ir - 0.5 * cdist;
                        unit = -cross(qp, rp);
                        unit = unit / length(unit);
                        V3 dv = p->rollingCondition
                                    ? of(pmi.vel) + cross(of(pmi.omega), po2cp) - (ivel + cross(iomega, ir * unit))
                                    : of(pmi.vel) - ivel;
                        xdem_device_force_constant c = force_constant(tcm, ir, 0, im, 0, cmp);
                        gpu_law(tcm, c, p->dt, 0.0, 0.0, rcon, cdist, iomega, dv, unit, Ft, Fn, M);
                        add_mag(id, Fn + Ft,
                                term_scale(c, cdist, 0.0,
                                           length(ivel) + length(iomega) * ir +
                                               length(of(pmi.vel) + cross(of(pmi.omega), po2cp)),
                                           Ft));
                        sum_force += Fn + Ft;
                        sum_moment += M;
                        if (rf) {
                            rf[pidx] += -(Fn + Ft);
                            rm[pidx] += -cross(po2cp, Fn + Ft);
                        }
                    }
                }
            }
}
}  // namespace

extern "C" {

void orc_polygon(const xdem_device_parameters* p, int32_t tcm, const xdem_device_polygon_info* dpi,
                 xdem_device_polygon_mass_info* dpmi, uint32_t nobj, int obj_reaction,
                 const double* pos4, const double* vel, const double* omega, double* force,
                 double* moment, const double* mass, const uint32_t* sorted,
                 const uint32_t* cstart, const uint32_t* cend,
                 const xdem_device_contact_property* cp, uint32_t np) {
    std::vector<V3> rf(nobj, mk(0, 0, 0)), rm(nobj, mk(0, 0, 0));
    for (uint32_t id = 0; id < np; ++id) {
        V3 sF, sM;
        std::vector<V3> pf(nobj, mk(0, 0, 0)), pm(nobj, mk(0, 0, 0));
        poly_one(p, tcm, id, dpi, dpmi, pos4, vel, omega, mass, sorted, cstart, cend, cp, np, sF,
                 sM, obj_reaction ? pf.data() : nullptr, obj_reaction ? pm.data() : nullptr);
        double* f = force + 3 * id;
        double* m = moment + 3 * id;
        f[0] += sF.x; f[1] += sF.y; f[2] += sF.z;
        m[0] += sM.x; m[1] += sM.y; m[2] += sM.z;
        if (obj_reaction)
            for (uint32_t o = 0; o < nobj; ++o) {
                rf[o] += pf[o];
                rm[o] += pm[o];
            }
    }
    if (obj_reaction)
        for (uint32_t o = 0; o < nobj; ++o) {
            dpmi[o].force = xdem_double3{rf[o].x, rf[o].y, rf[o].z};
            dpmi[o].moment = xdem_double3{rm[o].x, rm[o].y, rm[o].z};
        }
}

// vv_update_position_kernel (mphysics_cuda_impl.cuh:85-96): p += (dt v + half2dt a)
// velocity_verlet::updatePosition CPU (velocity_verlet.cpp:27-32): (p + dt v) + 0.5 dt^2 a (A10)
void orc_vv_position(const xdem_device_parameters* p, int semantics, double* pos4, const double* vel,
                     const double* acc, uint32_t np) {
    for (uint32_t id = 0; id < np; ++id) {
        V3 v = ld3(vel + 3 * id), a = ld3(acc + 3 * id);
        double* q = pos4 + 4 * id;
        if (semantics == ORC_HOST) {
            double sqt_dt = 0.5 * p->dt * p->dt;
            V3 np_ = ld3(q) + p->dt * v + sqt_dt * a;
            q[0] = np_.x; q[1] = np_.y; q[2] = np_.z;
        } else {
            V3 d = p->dt * v + p->half2dt * a;
            q[0] += d.x; q[1] += d.y; q[2] += d.z;
        }
    }
}

// vv_update_velocity_kernel (mphysics_cuda_impl.cuh:98-134) / CPU
// velocity_verlet.cpp:46-65 (two half kicks, A10).  Both reset F = m g, T = 0.
void orc_vv_velocity(const xdem_device_parameters* p, int semantics, double* vel, double* acc,
                     double* omega, double* alpha, double* force, double* moment,
                     const double* mass, const double* iner, uint32_t np) {
    V3 g = of(p->gravity);
    for (uint32_t id = 0; id < np; ++id) {
        double m = mass[id];
        V3 v = ld3(vel + 3 * id), av = ld3(omega + 3 * id);
        V3 F = ld3(force + 3 * id), T = ld3(moment + 3 * id);
        V3 a0 = ld3(acc + 3 * id), al0 = ld3(alpha + 3 * id);
        V3 a, in;
        if (semantics == ORC_HOST) {
            double inv_m = 1.0 / m, inv_i = 1.0 / iner[id];
            v += 0.5 * p->dt * a0;
            av += 0.5 * p->dt * al0;
            a = inv_m * F;
            in = inv_i * T;
            v += 0.5 * p->dt * a;
            av += 0.5 * p->dt * in;
        } else {
            a = (1.0 / m) * F;
            in = (1.0 / iner[id]) * T;
            v += 0.5 * p->dt * (a0 + a);
            av += 0.5 * p->dt * (al0 + in);
        }
        st3(force + 3 * id, m * g);
        st3(moment + 3 * id, mk(0.0, 0.0, 0.0));
        st3(vel + 3 * id, v);
        st3(omega + 3 * id, av);
        st3(acc + 3 * id, a);
        st3(alpha + 3 * id, in);
    }
}

// reduce6<double3,512> (mphysics_cuda_impl.cuh:776-835) launched as in
// reductionD3 (mphysics_cuda_def.cu:223-242): blocks of 512 threads, each
// block covering 1024 inputs, a shared-memory tree, block partials summed on
// the host in block order.
void orc_reduction_d3(const double* in3, uint32_t n, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (n == 0) return;
    const uint32_t B = 512;
    uint32_t nthreads = n < B ? n : B;
    uint32_t nblocks = (n + nthreads - 1) / nthreads;
    uint64_t gridSize = (uint64_t)B * 2 * nblocks;
    std::vector<V3> sdata(B);
    for (uint32_t b = 0; b < nblocks; ++b) {
        for (uint32_t tid = 0; tid < B; ++tid) {
            V3 s = mk(0.0, 0.0, 0.0);
            if (tid < nthreads) {
                uint64_t i = (uint64_t)b * (B * 2) + tid;
                while (i < n) {
                    s += ld3(in3 + 3 * i);
                    if (i + B < n) s += ld3(in3 + 3 * (i + B));
                    i += gridSize;
                }
            }
            sdata[tid] = s;
        }
        for (uint32_t s = 256; s >= 1; s >>= 1)
            for (uint32_t tid = 0; tid < s; ++tid) sdata[tid] = sdata[tid] + sdata[tid + s];
        out[0] += sdata[0].x;
        out[1] += sdata[0].y;
        out[2] += sdata[0].z;
    }
}

// numeric::utility::calculate_center_of_triangle — source/numeric_utility.cpp:20-45
void orc_circumcenter(const double P_[3], const double Q_[3], const double R_[3], double out[3]) {
    V3 P = ld3(P_), Q = ld3(Q_), R = ld3(R_);
    V3 V = Q - P, W = R - P;
    V3 N = cross(V, W);
    N = N / length(N);
    V3 M1 = 0.5 * (Q + P), M2 = 0.5 * (R + P);
    V3 D1 = cross(N, V), D2 = cross(N, W);
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
    st3(out, M1 + t * D1);
}

// updatePolygonObjectData_kernel — mphysics_cuda_impl.cuh:1898-1956
void orc_update_polygon(const xdem_device_polygon_mass_info* dpmi, const double* vList,
                        double* sphere4, xdem_device_polygon_info* dpi, uint32_t ntri) {
    for (uint32_t id = 0; id < ntri; ++id) {
        const xdem_device_polygon_mass_info& mi = dpmi[dpi[id].id];
        const xdem_double4& ep = mi.ep;
        V3 r0 = tfm1(ep), r1 = tfm2(ep), r2 = tfm3(ep);
        auto tg = [&](V3 v) {
            return mk(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
                      r2.x * v.x + r2.y * v.y + r2.z * v.z);
        };
        const double* s = vList + 9 * (size_t)id;
        V3 pos = of(mi.origin);
        V3 P = pos + tg(mk(s[0], s[1], s[2]));
        V3 Q = pos + tg(mk(s[3], s[4], s[5]));
        V3 R = pos + tg(mk(s[6], s[7], s[8]));
        double Pa[3] = {P.x, P.y, P.z}, Qa[3] = {Q.x, Q.y, Q.z}, Ra[3] = {R.x, R.y, R.z}, c[3];
        orc_circumcenter(Pa, Qa, Ra, c);
        sphere4[4 * id + 0] = c[0];
        sphere4[4 * id + 1] = c[1];
        sphere4[4 * id + 2] = c[2];
        dpi[id].P = xdem_double3{P.x, P.y, P.z};
        dpi[id].Q = xdem_double3{Q.x, Q.y, Q.z};
        dpi[id].R = xdem_double3{R.x, R.y, R.z};
    }
}

// Pose after tau seconds of motion: origin0 + v tau, q(omega, tau) * ep0
// (Hamilton product, e0 first), so A(ep) = A(q) A(ep0).
void orc_motion_pose(const xdem_double3* o0, const xdem_double4* ep0, const xdem_motion* m,
                     double tau, xdem_double3* o, xdem_double4* ep) {
    *o = xdem_double3{o0->x + tau * m->velocity.x, o0->y + tau * m->velocity.y,
                      o0->z + tau * m->velocity.z};
    const xdem_double3& w = m->omega;
    double wn = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    if (!(wn > 0)) {
        *ep = *ep0;
        return;
    }
    double h = 0.5 * (wn * tau);
    double c = std::cos(h), sn = std::sin(h);
    double q0 = c, q1 = sn * (w.x / wn), q2 = sn * (w.y / wn), q3 = sn * (w.z / wn);
    const xdem_double4& a = *ep0;
    xdem_double4 r;
    r.x = q0 * a.x - (q1 * a.y + q2 * a.z + q3 * a.w);
    r.y = q0 * a.y + a.x * q1 + (q2 * a.w - q3 * a.z);
    r.z = q0 * a.z + a.x * q2 + (q3 * a.y - q1 * a.w);
    r.w = q0 * a.w + a.x * q3 + (q1 * a.z - q2 * a.y);
    *ep = r;
}

// toLocal (mphysics_cuda_impl.cuh:640-645) of each vertex relative to its
// object's origin.
void orc_local_vertices(const xdem_device_polygon_info* tris, uint32_t ntri,
                        const xdem_device_polygon_mass_info* objs, double* vlist) {
    for (uint32_t t = 0; t < ntri; ++t) {
        const xdem_device_polygon_mass_info& o = objs[tris[t].id];
        V3 r0 = tfm1(o.ep), r1 = tfm2(o.ep), r2 = tfm3(o.ep);
        const xdem_double3* v[3] = {&tris[t].P, &tris[t].Q, &tris[t].R};
        for (int j = 0; j < 3; ++j) {
            V3 d = mk(v[j]->x - o.origin.x, v[j]->y - o.origin.y, v[j]->z - o.origin.z);
            st3(vlist + 9 * (size_t)t + 3 * j, to_local(r0, r1, r2, d));
        }
    }
}

}  // extern "C"

namespace {

// Geometry motion of step w->step (xDynamicsSolver.cpp:170-184): time
// ct = step*dt, every motion whose start has come moves one step; moved planes
// are rewritten (contact_particles_plane.cpp:33-49), moved objects get their
// pose and velocity and then every triangle is re-posed from its local
// coordinates (updatePolygonObjectData_kernel).
void apply_motions(orc_world* w) {
    const double dt = w->prm.dt;
    const double t = (double)w->step * dt;
    bool repose = false;
    for (uint32_t i = 0; i < w->nmotion; ++i) {
        orc_motion& mo = w->motions[i];
        const xdem_motion& m = mo.m;
        bool active = t >= m.start_time && (!(m.end_time > m.start_time) || t <= m.end_time);
        if (active) mo.n_act++;
        const double tau = (double)mo.n_act * dt;
        if (mo.kind == XDEM_OBJECT_PLANE) {
            if (!active) continue;
            xdem_device_plane_info pl = w->planes0[mo.index];
            for (xdem_double3* q : {&pl.xw, &pl.w2, &pl.w3, &pl.w4})
                *q = xdem_double3{q->x + tau * m.velocity.x, q->y + tau * m.velocity.y,
                                  q->z + tau * m.velocity.z};
            w->planes[mo.index] = pl;
        } else {
            if (!active && (mo.active != 0) == active) continue;
            const xdem_device_polygon_mass_info& b = w->obj0[mo.index];
            xdem_device_polygon_mass_info o = b;
            orc_motion_pose(&b.origin, &b.ep, &m, tau, &o.origin, &o.ep);
            o.vel = active ? m.velocity : xdem_double3{0, 0, 0};
            o.omega = active ? m.omega : xdem_double3{0, 0, 0};
            o.force = xdem_double3{0, 0, 0};
            o.moment = xdem_double3{0, 0, 0};
            w->obj[mo.index] = o;
            repose = true;
        }
        mo.active = active ? 1 : 0;
    }
    if (repose) orc_update_polygon(w->obj, w->vlist, w->tri_sphere, w->tri, w->ntri);
}

// Walls of one particle in the CPU path (contactManager.cpp:228-234): a lone
// plane assigns F/M (contact_particles_plane.cpp:72,233 -> contact.cpp:43,45,
// A2); a cube plane accumulates through locals (contact_particles_cube.cpp:27-33).
void host_walls_one(const orc_world* w, uint32_t id, V3& F, V3& M) {
    const double* ip = w->pos + 4 * id;
    double r = ip[3], m = w->mass[id];
    V3 pos = ld4(ip), vel = ld3(w->vel + 3 * id);
    for (uint32_t k = 0; k < w->nwall; ++k) {
        const orc_wall& wl = w->walls[k];
        if (wl.kind != ORC_WALL_PLANE) continue;  // CPU cylinder class is commented out
        const xdem_device_plane_info* pe = &w->planes[wl.index];
        V3 dp = pos - of(pe->xw);
        double wp[3] = {dot(dp, of(pe->u1)), dot(dp, of(pe->u2)), dot(dp, of(pe->uw))};
        double u[3] = {0, 0, 0}, xp[3] = {pos.x, pos.y, pos.z};
        double cdist = orc_plane_detect(pe, ORC_HOST, xp, wp, u, r);
        if (!(cdist > 0)) continue;
        double rcon = r - 0.5 * cdist;
        V3 un = ld3(u);
        V3 cpt = rcon * un;
        V3 dv = -(vel);  // A3
        xdem_device_force_constant c = force_constant(1, r, 0.0, m, 0.0, wl.cp);
        double coh_e = ((1.0 - wl.cp.pri * wl.cp.pri) / wl.cp.Ei) + ((1.0 - wl.cp.prj * wl.cp.prj) / wl.cp.Ej);
        double fsn = (-c.kn * std::pow(cdist, 1.5));
        double fca = host_cohesion(wl.cp.coh, r, coh_e, fsn);
        if (wl.cube_member) {
            V3 mf = mk(0, 0, 0), mm = mk(0, 0, 0);
            host_dhs(c, w->prm.dt, fca, cdist, cpt, dv, un, mf, mm);
            F += mf;
            M += mm;
        } else {
            host_dhs(c, w->prm.dt, fca, cdist, cpt, dv, un, F, M);
        }
    }
}

void detection(orc_world* w) {
    uint32_t np = w->np;
    int sem = w->semantics;
    orc_hash_and_index(&w->prm, sem, w->pos, np, 0, w->hash, w->index);
    if (w->ntri) orc_hash_and_index(&w->prm, sem, w->tri_sphere, w->ntri, np, w->hash, w->index);
    sort_cells(w->hash, w->index, np + w->ntri, w->prm.ncell, w->cstart, w->cend, w->sorted);
}

void step_gpu(orc_world* w, int nthreads) {
    const xdem_device_parameters* p = &w->prm;
    uint32_t np = w->np;
    w->step++;
    if (w->nmotion) apply_motions(w);
    orc_vv_position(p, ORC_GPU, w->pos, w->vel, w->acc, np);
    detection(w);
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        uint32_t id = (uint32_t)i;
        V3 sF, sM;
        uint64_t dummy = 0;
        p2p_one(p, w->tcm, ORC_GPU, id, w->pos, w->vel, w->omega, w->mass, w->sorted, w->cstart,
                w->cend, w->cp_pp, np, sF, sM, nullptr, 0, &dummy);
        double* f = w->force + 3 * id;
        double* m = w->moment + 3 * id;
        f[0] += sF.x; f[1] += sF.y; f[2] += sF.z;
        m[0] += sM.x; m[1] += sM.y; m[2] += sM.z;
    }
    for (uint32_t k = 0; k < w->nwall; ++k) {
        const orc_wall& wl = w->walls[k];
        if (wl.kind == ORC_WALL_PLANE)
            orc_plane(p, w->tcm, &w->planes[wl.index], w->pos, w->vel, w->omega, w->force,
                      w->moment, w->mass, np, &wl.cp);
        else if (wl.kind == ORC_WALL_CYLINDER)
            orc_cylinder(p, w->tcm, &w->cylinders[wl.index], w->pos, w->vel, w->omega, w->force,
                         w->moment, w->mass, np, &wl.cp, nullptr, nullptr, nullptr);
        else
            orc_drum(p, w->tcm, &w->drums[wl.index], w->pos, w->vel, w->omega, w->force,
                     w->moment, w->mass, np, &wl.cp, nullptr, nullptr);
    }
    if (w->ntri)
        orc_polygon(p, w->tcm, w->tri, w->obj, w->nobj, 1, w->pos, w->vel, w->omega, w->force,
                    w->moment, w->mass, w->sorted, w->cstart, w->cend, w->cp_obj, np);
    orc_vv_velocity(p, ORC_GPU, w->vel, w->acc, w->omega, w->alpha, w->force, w->moment, w->mass,
                    w->iner, np);
}

// contactManager::hostCollision — contactManager.cpp:208-277
void step_host(orc_world* w, int nthreads) {
    const xdem_device_parameters* p = &w->prm;
    uint32_t np = w->np;
    w->step++;
    if (w->nmotion) apply_motions(w);
    orc_vv_position(p, ORC_HOST, w->pos, w->vel, w->acc, np);
    detection(w);
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        uint32_t id = (uint32_t)i;
        V3 F = mk(0, 0, 0), M = mk(0, 0, 0);
        host_walls_one(w, id, F, M);
        V3 sF, sM;
        uint64_t dummy = 0;
        p2p_one(p, 1, ORC_HOST, id, w->pos, w->vel, w->omega, w->mass, w->sorted, w->cstart,
                w->cend, w->cp_pp, np, sF, sM, nullptr, 0, &dummy);
        F += sF;
        M += sM;
        if (w->ntri) {
            V3 pF, pM;
            poly_one(p, 1, id, w->tri, w->obj, w->pos, w->vel, w->omega, w->mass, w->sorted,
                     w->cstart, w->cend, w->cp_obj, np, pF, pM, nullptr, nullptr);
            F += pF;
            M += pM;
        }
        double* f = w->force + 3 * id;
        double* m = w->moment + 3 * id;
        f[0] += F.x; f[1] += F.y; f[2] += F.z;
        m[0] += M.x; m[1] += M.y; m[2] += M.z;
    }
    orc_vv_velocity(p, ORC_HOST, w->vel, w->acc, w->omega, w->alpha, w->force, w->moment, w->mass,
                    w->iner, np);
}

// The release of the coming step (w->step = k - 1 before step k): the next
// group joins the simulated prefix when (k - 1) % release_every == 0.
void release_step(orc_world* w) {
    if (w->release_every && w->step % w->release_every == 0 && w->np < w->np_total) {
        uint32_t add = w->release_group;
        if (w->nrelease_groups)
            add = w->release_next < w->nrelease_groups ? w->release_groups[w->release_next++] : 0;
        w->np = std::min(w->np_total, w->np + add);
    }
}

}  // namespace

extern "C" {

void orc_step(orc_world* w, uint32_t nsteps) {
    for (uint32_t s = 0; s < nsteps; ++s) {
        release_step(w);
        if (w->semantics == ORC_HOST)
            step_host(w, 1);
        else
            step_gpu(w, 1);
    }
}

void orc_step_omp(orc_world* w, uint32_t nsteps, int32_t nthreads) {
    for (uint32_t s = 0; s < nsteps; ++s) {
        release_step(w);
        if (w->semantics == ORC_HOST)
            step_host(w, nthreads);
        else
            step_gpu(w, nthreads);
    }
}

}  // extern "C"

namespace {

// Tangential-spring DHS (no reference counterpart; the law is stated in
// xdem_oracle.h).  xi is in/out.
void hist_law(const xdem_device_force_constant& c, double dt, double ir, double cdist, V3 dv,
              V3 unit, V3& xi, V3& Ft, V3& Fn, V3& M, double fca = 0.0) {
    double fsn = -c.kn * std::pow(cdist, 1.5);
    double fdn = c.vn * dot(dv, unit);
    Fn = (fsn + fca + fdn) * unit;
    V3 e = dv - dot(dv, unit) * unit;
    xi = xi - dot(xi, unit) * unit;
    xi = xi + dt * e;
    Ft = c.ks * xi + c.vs * e;
    double ft = length(Ft), fmax = c.mu * length(Fn);
    if (ft > fmax) {
        // sliding: Ft at the Coulomb cap, the spring reset to the value that
        // reproduces it (ks xi + vs e = Ft)
        Ft = (fmax / ft) * Ft;
        xi = c.ks > 0 ? (1.0 / c.ks) * (Ft - c.vs * e) : mk(0, 0, 0);
    }
    M = cross(ir * unit, Ft);
}

struct HistCursor {
    const orc_hist* in;
    orc_hist* out;
    uint32_t id;
    uint32_t nout = 0;
    V3 lookup(uint32_t key) const {
        uint32_t H = in->slots, n = in->count[id];
        for (uint32_t k = 0; k < n; ++k)
            if (in->key[(size_t)id * H + k] == key) {
                const double* x = in->xi + 3 * ((size_t)id * H + k);
                return mk(x[0], x[1], x[2]);
            }
        return mk(0, 0, 0);
    }
    void store(uint32_t key, V3 xi, uint64_t* overflow) {
        uint32_t H = out->slots;
        if (nout < H) {
            out->key[(size_t)id * H + nout] = key;
            st3(out->xi + 3 * ((size_t)id * H + nout), xi);
            ++nout;
        } else {
            ++*overflow;
        }
    }
};

// Contacts of particle id with history.  F and Mo enter with the values the
// contacts are added to (the kernels' order: F += particle sum, then F += each
// plane's force, then each drum contact's).
void hist_one(const orc_world* w, const orc_hist* in, orc_hist* out, uint32_t id, V3& F, V3& Mo,
              uint64_t* overflow) {
    const xdem_device_parameters* p = &w->prm;
    const double* pos4 = w->pos;
    V3 ipos = ld4(pos4 + 4 * id);
    double ir = pos4[4 * id + 3];
    V3 ivel = ld3(w->vel + 3 * id), iomega = ld3(w->omega + 3 * id);
    double im = w->mass[id];
    int32_t g[3];
    orc_cell_of(p, ORC_GPU, ipos.x, ipos.y, ipos.z, g);
    HistCursor hc{in, out, id};
    V3 sumF = mk(0, 0, 0), sumM = mk(0, 0, 0);
    // particle contacts: calculate_p2p_kernel traversal (mphysics_cuda_impl.cuh:430-482)
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                int32_t nb[3] = {g[0] + x, g[1] + y, g[2] + z};
                uint32_t h = orc_hash_of(p, nb);
                uint32_t s = w->cstart[h];
                if (s == 0xFFFFFFFFu) continue;
                uint32_t e = w->cend[h];
                for (uint32_t j = s; j < e; ++j) {
                    uint32_t k = w->sorted[j];
                    if (id == k || k >= w->np) continue;
                    V3 jpos = ld4(pos4 + 4 * k);
                    double jr = pos4[4 * k + 3];
                    V3 rp = mk(jpos.x - ipos.x, jpos.y - ipos.y, jpos.z - ipos.z);
                    double dist = length(rp);
                    double cdist = (ir + jr) - dist;
                    if (!(cdist > 0)) continue;
                    V3 unit = rp / dist;
                    V3 jvel = ld3(w->vel + 3 * k), jomega = ld3(w->omega + 3 * k);
                    V3 rv = p->rollingCondition
                                ? jvel + cross(jomega, -jr * unit) - (ivel + cross(iomega, ir * unit))
                                : (jvel - ivel);
                    xdem_device_force_constant c = force_constant(1, ir, jr, im, w->mass[k], w->cp_pp);
                    // cohesionForce on particle pairs, as the DHS kernel adds it
                    const xdem_device_contact_property& cp = w->cp_pp;
                    double fca = cohesion_force(ir, jr, cp.Ei, cp.Ej, cp.pri, cp.prj, cp.coh,
                                                -c.kn * std::pow(cdist, 1.5));
                    V3 xi = hc.lookup(k), Ft, Fn, M;
                    hist_law(c, p->dt, ir, cdist, rv, unit, xi, Ft, Fn, M, fca);
                    hc.store(k, xi, overflow);
                    add_mag(id, Fn + Ft,
                            term_scale(c, cdist, fca,
                                       length(ivel) + length(iomega) * ir + length(jvel) +
                                           length(jomega) * jr,
                                       Ft));
                    sumF += Fn + Ft;
                    sumM += M;
                }
            }
    F += sumF;
    Mo += sumM;
    // planes in application order (plane_contact_force_kernel, :573-621; ir = 0)
    for (uint32_t q = 0; q < w->nwall; ++q) {
        const orc_wall& wl = w->walls[q];
        if (wl.kind != ORC_WALL_PLANE) continue;
        const xdem_device_plane_info* pl = &w->planes[wl.index];
        V3 dp = ipos - of(pl->xw);
        double unit[3] = {0, 0, 0};
        double wp[3] = {dot(dp, of(pl->u1)), dot(dp, of(pl->u2)), dot(dp, of(pl->uw))};
        double xp[3] = {ipos.x, ipos.y, ipos.z};
        double cdist = orc_plane_detect(pl, ORC_GPU, xp, wp, unit, ir);
        if (!(cdist > 0)) continue;
        V3 un = ld3(unit);
        V3 dv = p->rollingCondition ? -(ivel + cross(iomega, ir * un)) : -ivel;
        xdem_device_force_constant c = force_constant(1, ir, 0.0, im, 0.0, wl.cp);
        uint32_t key = ORC_WALL_KEY | q;
        V3 xi = hc.lookup(key), Ft, Fn, M;
        hist_law(c, p->dt, 0.0, cdist, dv, un, xi, Ft, Fn, M);
        hc.store(key, xi, overflow);
        add_mag(id, Fn + Ft, term_scale(c, cdist, 0.0, length(ivel) + length(iomega) * ir, Ft));
        F += Fn + Ft;
    }
    // drums in order: orc_drum's shell and cap contacts (the cylinder kernel's
    // evaluation, ir = 0) with a spring per surface
    for (uint32_t q = 0; q < w->nwall; ++q) {
        const orc_wall& wl = w->walls[q];
        if (wl.kind != ORC_WALL_DRUM) continue;
        const xdem_drum_info* dr = &w->drums[wl.index];
        V3 ax = of(dr->axis), o = of(dr->origin);
        V3 d = ipos - o;
        double a = dot(d, ax);
        if (!(std::fabs(a) <= dr->half_length)) continue;
        V3 qv = d - a * ax;
        double rho = std::sqrt(dot(qv, qv));
        auto contact = [&](V3 un, V3 cpt, double cdist, uint32_t side) {
            V3 wv = dr->omega * cross(ax, cpt - o);
            V3 dv = wv - (ivel + cross(iomega, ir * un));
            xdem_device_force_constant c = force_constant(1, ir, 0.0, im, 0.0, wl.cp);
            uint32_t key = ORC_DRUM_KEY | (wl.index << 2) | side;
            V3 xi = hc.lookup(key), Ft, Fn, M;
            hist_law(c, p->dt, 0.0, cdist, dv, un, xi, Ft, Fn, M);
            hc.store(key, xi, overflow);
            add_mag(id, Fn + Ft,
                    term_scale(c, cdist, 0.0,
                               length(ivel) + length(iomega) * ir + length(wv), Ft));
            F += Fn + Ft;
        };
        double cs = ir + rho - dr->radius;
        if (cs > 0 && rho > 0) {
            V3 u = qv / rho;
            contact(u, o + a * ax + dr->radius * u, cs, 0u);
        }
        double cc = ir - (dr->half_length - std::fabs(a));
        if (dr->caps && cc > 0 && rho < dr->radius) {
            double sg = a >= 0 ? 1.0 : -1.0;
            contact(sg * ax, o + (sg * dr->half_length) * ax + qv, cc, a >= 0 ? 1u : 2u);
        }
    }
    out->count[id] = hc.nout;
}

}  // namespace

extern "C" {

void orc_detect(orc_world* w) { detection(w); }

void orc_hist_contacts(const orc_world* w, const orc_hist* in, orc_hist* out, double* force,
                       double* moment) {
    uint64_t ov = 0;
    for (uint32_t id = 0; id < w->np; ++id) {
        V3 F = mk(0, 0, 0), M = mk(0, 0, 0);
        hist_one(w, in, out, id, F, M, &ov);
        st3(force + 3 * id, F);
        st3(moment + 3 * id, M);
    }
    out->overflow = ov;
}

void orc_hist_step(orc_world* w, const orc_hist* in, orc_hist* out) {
    const xdem_device_parameters* p = &w->prm;
    uint32_t np = w->np;
    orc_vv_position(p, ORC_GPU, w->pos, w->vel, w->acc, np);
    detection(w);
    uint64_t ov = 0;
    for (uint32_t id = 0; id < np; ++id) {
        // force holds m g (the previous velocity update's reset, :128)
        V3 F = ld3(w->force + 3 * id), M = ld3(w->moment + 3 * id);
        hist_one(w, in, out, id, F, M, &ov);
        st3(w->force + 3 * id, F);
        st3(w->moment + 3 * id, M);
    }
    out->overflow = ov;
    orc_vv_velocity(p, ORC_GPU, w->vel, w->acc, w->omega, w->alpha, w->force, w->moment, w->mass,
                    w->iner, np);
}

}  // extern "C"
