// xdem_host.hip — host-side counterparts of the reference's algorithm objects,
// driving the plain FP64 kernels through their launchers:
//   grid (xdem_grid)  <- grid_base / neighborhood_cell (include/grid_base.h:7-64,
//                        source/grid_base.cpp:79-155, source/neighborhood_cell.cpp:91-107)
//   sim  (xdem_sim)   <- dem_simulation::initialize / oneStepAnalysis
//                        (source/simulation/dem_simulation.cpp:204-325, 447-469),
//                        velocity_verlet (source/velocity_verlet.cpp:18-67),
//                        contactManager::deviceCollision (source/contactManager.cpp:156-185)
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/xdem_io.h"
#include "xdem_common.hpp"
#include "xdem_f64.hpp"
#include "xdem_io.hpp"

using namespace xdem;

struct xdem_grid {
    xdem_double3 wo{-1.0, -1.0, -1.0};
    double cs = 0.0;
    xdem_uint3 gs{128, 128, 128};
    uint32_t nse = 0;  // number of sorting elements
    uint32_t ng = 0;   // number of cells
    uint32_t* d_cell_id = nullptr;
    uint32_t* d_body_id = nullptr;
    uint32_t* d_sorted_id = nullptr;
    uint32_t* d_cell_start = nullptr;
    uint32_t* d_cell_end = nullptr;

    KParams kparams() const {
        xdem_device_parameters p{};
        p.grid_size = gs;
        p.ncell = gs.x * gs.y * gs.z;
        p.cell_size = cs;
        p.world_origin = wo;
        return params_from(p);
    }
    void release() {
        // grid_base::clear (grid_base.cpp:59-77), without its d_cell_start/d_cell_end mix-up.
        for (uint32_t** p : {&d_cell_id, &d_body_id, &d_sorted_id, &d_cell_start, &d_cell_end}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
    ~xdem_grid() { release(); }
};

static int32_t grid_alloc(xdem_grid* g, uint32_t n) {
    XDEM_REQUIRE(g->gs.x && g->gs.y && g->gs.z, "grid size not set");
    XDEM_REQUIRE(g->cs > 0, "cell size not set");
    xdem_device_parameters p{};
    p.grid_size = g->gs;
    p.ncell = g->gs.x * g->gs.y * g->gs.z;
    p.cell_size = g->cs;
    XDEM_TRY(validate_params(p));
    g->release();
    g->ng = p.ncell;
    g->nse = n;
    size_t nn = std::max<uint32_t>(n, 1);
    XDEM_HIP(hipMalloc(&g->d_cell_id, sizeof(uint32_t) * nn));
    XDEM_HIP(hipMalloc(&g->d_body_id, sizeof(uint32_t) * nn));
    XDEM_HIP(hipMalloc(&g->d_sorted_id, sizeof(uint32_t) * nn));
    XDEM_HIP(hipMalloc(&g->d_cell_start, sizeof(uint32_t) * g->ng));
    XDEM_HIP(hipMalloc(&g->d_cell_end, sizeof(uint32_t) * g->ng));
    return XDEM_OK;
}

// neighborhood_cell::detection (GPU branch): hash particles (K3), hash triangle
// circumspheres at sid = np (K4), then sort + cell ranges (K5-K7).
static int32_t grid_detect(xdem_grid* g, const double* pos, const double* spos, uint32_t np,
                           uint32_t snp, hipStream_t s, PhaseTimer* tm = nullptr) {
    PhaseTimer dummy;
    PhaseTimer& T = tm ? *tm : dummy;
    XDEM_REQUIRE(g->d_cell_id, "grid not initialized");
    XDEM_REQUIRE(np + snp <= g->nse, "np + snp (%u) exceeds the grid capacity %u", np + snp,
                 g->nse);
    KParams k = g->kparams();
    XDEM_TRY(T.begin(1, s));
    XDEM_TRY(launch_hash(k, g->d_cell_id, g->d_body_id, pos, 0, np, s));
    if (snp && spos) XDEM_TRY(launch_hash(k, g->d_cell_id, g->d_body_id, spos, np, snp, s));
    XDEM_TRY(T.end(s));
    uint32_t n = np + (spos ? snp : 0);
    XDEM_TRY(T.begin(2, s));
    XDEM_TRY(sort_pairs(g->d_cell_id, g->d_body_id, n, bits_for(g->ng), s));
    XDEM_TRY(T.end(s));
    XDEM_TRY(T.begin(3, s));
    XDEM_TRY(fill_u32(g->d_cell_start, 0xFFFFFFFFu, g->ng, s));
    XDEM_TRY(fill_u32(g->d_cell_end, 0u, g->ng, s));
    XDEM_TRY(cell_ranges(g->d_cell_id, g->d_body_id, g->d_cell_start, g->d_cell_end,
                         g->d_sorted_id, n, s));
    XDEM_TRY(T.end(s));
    return XDEM_OK;
}

extern "C" {

int32_t xdem_grid_create(xdem_grid** out) {
    XDEM_REQUIRE(out, "null out");
    *out = new xdem_grid();
    return XDEM_OK;
}
int32_t xdem_grid_destroy(xdem_grid* g) {
    delete g;
    return XDEM_OK;
}
int32_t xdem_grid_set_world_origin(xdem_grid* g, double x, double y, double z) {
    XDEM_REQUIRE(g, "null grid");
    g->wo = xdem_double3{x, y, z};
    return XDEM_OK;
}
int32_t xdem_grid_set_cell_size(xdem_grid* g, double cs) {
    XDEM_REQUIRE(g && cs > 0, "invalid cell size");
    g->cs = cs;
    return XDEM_OK;
}
int32_t xdem_grid_set_grid_size(xdem_grid* g, uint32_t gx, uint32_t gy, uint32_t gz) {
    XDEM_REQUIRE(g, "null grid");
    g->gs = xdem_uint3{gx, gy, gz};
    return XDEM_OK;
}
int32_t xdem_grid_initialize(xdem_grid* g, uint32_t n) {
    XDEM_REQUIRE(g, "null grid");
    return grid_alloc(g, n);
}
int32_t xdem_grid_detection(xdem_grid* g, const double* pos, const double* spos, uint32_t np,
                            uint32_t snp) {
    XDEM_REQUIRE(g && (np == 0 || pos), "null argument");
    return grid_detect(g, pos, spos, np, snp, stream());
}
uint32_t xdem_grid_ncell(const xdem_grid* g) { return g ? g->ng : 0; }
uint32_t* xdem_grid_sorted_id(xdem_grid* g) { return g ? g->d_sorted_id : nullptr; }
uint32_t* xdem_grid_cell_start(xdem_grid* g) { return g ? g->d_cell_start : nullptr; }
uint32_t* xdem_grid_cell_end(xdem_grid* g) { return g ? g->d_cell_end : nullptr; }
uint32_t* xdem_grid_cell_id(xdem_grid* g) { return g ? g->d_cell_id : nullptr; }
uint32_t* xdem_grid_body_id(xdem_grid* g) { return g ? g->d_body_id : nullptr; }

// grid_base::getCellNumber (grid_base.cpp:139-146; the GPU floor form, no abs()).
int32_t xdem_grid_get_cell_number(const xdem_grid* g, double x, double y, double z,
                                  int32_t out[3]) {
    XDEM_REQUIRE(g && out && g->cs > 0, "invalid grid");
    out[0] = (int32_t)std::floor((x - g->wo.x) / g->cs);
    out[1] = (int32_t)std::floor((y - g->wo.y) / g->cs);
    out[2] = (int32_t)std::floor((z - g->wo.z) / g->cs);
    return XDEM_OK;
}
// grid_base::getHash (grid_base.cpp:148-155)
uint32_t xdem_grid_get_hash(const xdem_grid* g, const int32_t c3[3]) {
    uint32_t x = (uint32_t)c3[0] & (g->gs.x - 1);
    uint32_t y = (uint32_t)c3[1] & (g->gs.y - 1);
    uint32_t z = (uint32_t)c3[2] & (g->gs.z - 1);
    return (z * g->gs.y) * g->gs.x + (y * g->gs.x) + x;
}

}  // extern "C"

// ================================================================== sim
namespace {

struct WallOp {
    int kind;        // 0 = run of planes, 1 = cylinder, 2 = drum
    uint32_t first;  // into planes / cylinders / drums
    uint32_t count;
};

// A prescribed motion (xdem_motion) of plane or polygon object `index`;
// n_act counts the steps on which it was active.
struct Motion {
    int32_t kind;
    uint32_t index;
    xdem_motion m;
    uint64_t n_act;
    bool active;
};

// makeTFM_1..3 (mphysics_cuda_impl.cuh:624-638): rows of A(ep), e0 = ep.x.
void ep_rows(const xdem_double4& ep, d3& r0, d3& r1, d3& r2) {
    r0 = mk3(2.0 * (ep.x * ep.x + ep.y * ep.y - 0.5), 2.0 * (ep.y * ep.z - ep.x * ep.w),
             2.0 * (ep.y * ep.w + ep.x * ep.z));
    r1 = mk3(2.0 * (ep.y * ep.z + ep.x * ep.w), 2.0 * (ep.x * ep.x + ep.z * ep.z - 0.5),
             2.0 * (ep.z * ep.w - ep.x * ep.y));
    r2 = mk3(2.0 * (ep.y * ep.w - ep.x * ep.z), 2.0 * (ep.z * ep.w + ep.x * ep.y),
             2.0 * (ep.x * ep.x + ep.w * ep.w - 0.5));
}

// Pose of a moving object after tau seconds of motion: origin0 + v tau and
// q(omega, tau) * ep0 (Hamilton product, scalar first).
void motion_pose(const xdem_double3& o0, const xdem_double4& ep0, const xdem_motion& m,
                 double tau, xdem_double3& o, xdem_double4& ep) {
    o = xdem_double3{o0.x + tau * m.velocity.x, o0.y + tau * m.velocity.y,
                     o0.z + tau * m.velocity.z};
    const xdem_double3& w = m.omega;
    double wn = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    if (!(wn > 0)) {
        ep = ep0;
        return;
    }
    double h = 0.5 * (wn * tau);
    double c = std::cos(h), sn = std::sin(h);
    double q0 = c, q1 = sn * (w.x / wn), q2 = sn * (w.y / wn), q3 = sn * (w.z / wn);
    const xdem_double4& p = ep0;
    ep.x = q0 * p.x - (q1 * p.y + q2 * p.z + q3 * p.w);
    ep.y = q0 * p.y + p.x * q1 + (q2 * p.w - q3 * p.z);
    ep.z = q0 * p.z + p.x * q2 + (q3 * p.y - q1 * p.w);
    ep.w = q0 * p.w + p.x * q3 + (q1 * p.z - q2 * p.y);
}

template <class T>
int32_t upload(T** d, const std::vector<T>& h) {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    if (h.empty()) return XDEM_OK;
    XDEM_HIP(hipMalloc(d, sizeof(T) * h.size()));
    XDEM_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return XDEM_OK;
}

// device_contact_property as contact::cudaMemoryAlloc builds it (contact.cpp:107-116).
xdem_device_contact_property make_cp(const xdem_material& mi, const xdem_contact_pair_desc& c) {
    auto shear = [](const xdem_material& m) {
        return m.shear > 0 ? m.shear : m.youngs / (2.0 * (1.0 + m.poisson));
    };
    const xdem_material& mj = c.wall;
    return xdem_device_contact_property{mi.youngs, mj.youngs, mi.poisson, mj.poisson,
                                        shear(mi), shear(mj), c.restitution, c.friction,
                                        c.rolling_friction, c.cohesion, c.stiffness_ratio};
}

}  // namespace

struct xdem_sim {
    xdem_sim_desc desc{};
    xdem_device_parameters prm{};
    uint32_t np = 0;
    double* pos = nullptr;
    double* vel = nullptr;
    double* omega = nullptr;
    double* acc = nullptr;
    double* alpha = nullptr;
    double* force = nullptr;
    double* moment = nullptr;
    double* mass = nullptr;
    double* iner = nullptr;
    double max_radius = 0.0;
    xdem_grid grid;
    xdem_device_contact_property* d_cp_pp = nullptr;
    std::vector<xdem_device_plane_info> h_planes;
    std::vector<xdem_device_contact_property> h_plane_cp;
    std::vector<xdem_device_cylinder_info> h_cyl;
    std::vector<xdem_device_contact_property> h_cyl_cp;
    std::vector<xdem_drum_info> h_drum;
    std::vector<xdem_device_contact_property> h_drum_cp;
    std::vector<WallOp> ops;
    xdem_drum_info* d_drum = nullptr;
    xdem_device_contact_property* d_drum_cp = nullptr;
    xdem_device_plane_info* d_planes = nullptr;
    xdem_device_contact_property* d_plane_cp = nullptr;
    xdem_device_cylinder_info* d_cyl = nullptr;
    xdem_device_contact_property* d_cyl_cp = nullptr;
    // triangles
    uint32_t ntri = 0, nobj = 0;
    xdem_device_polygon_info* d_tri = nullptr;
    double* d_sph = nullptr;
    xdem_device_polygon_mass_info* d_obj = nullptr;
    xdem_device_contact_property* d_obj_cp = nullptr;
    std::vector<xdem_device_polygon_mass_info> h_obj;  // poses at set time
    std::vector<double> h_vlist;                         // 9 local coordinates per triangle
    double* d_vlist = nullptr;
    std::vector<Motion> motions;
    uint64_t step = 0;
    // real-time release: the first np_act particles are simulated; every
    // release_every steps (from step 1) the next group joins: release_group
    // particles, or the next entry of release_groups when that list is set
    // (0: all, always)
    uint32_t release_every = 0, release_group = 0, np_act = 0, release_next = 0;
    std::vector<uint32_t> release_groups;
    uint32_t active() const { return release_every ? np_act : np; }
    uint32_t next_group() {
        if (release_groups.empty()) return release_group;
        return release_next < release_groups.size() ? release_groups[release_next++] : 0;
    }
    bool dirty = true;
    PhaseTimer timer;
    float4* d_pfs = nullptr;  // sorted FP32 screening copy for the particle pass
    uint32_t pfs_cap = 0;

    ~xdem_sim() {
        for (void* p : {(void*)pos, (void*)vel, (void*)omega, (void*)acc, (void*)alpha,
                        (void*)force, (void*)moment, (void*)mass, (void*)iner, (void*)d_cp_pp,
                        (void*)d_planes, (void*)d_plane_cp, (void*)d_cyl, (void*)d_cyl_cp,
                        (void*)d_tri, (void*)d_sph, (void*)d_obj, (void*)d_obj_cp,
                        (void*)d_drum, (void*)d_drum_cp, (void*)d_vlist, (void*)d_pfs})
            if (p) (void)hipFree(p);
    }
    KParams kparams() const { return params_from(prm); }
};

static int32_t sim_sync_walls(xdem_sim* s) {
    if (!s->dirty) return XDEM_OK;
    XDEM_TRY(upload(&s->d_planes, s->h_planes));
    XDEM_TRY(upload(&s->d_plane_cp, s->h_plane_cp));
    XDEM_TRY(upload(&s->d_cyl, s->h_cyl));
    XDEM_TRY(upload(&s->d_cyl_cp, s->h_cyl_cp));
    XDEM_TRY(upload(&s->d_drum, s->h_drum));
    XDEM_TRY(upload(&s->d_drum_cp, s->h_drum_cp));
    s->dirty = false;
    return XDEM_OK;
}

static int32_t sim_regrid(xdem_sim* s) {
    double cs = s->desc.cell_size > 0 ? s->desc.cell_size : 2.0 * s->max_radius;
    s->prm.cell_size = cs;
    s->grid.wo = s->prm.world_origin;
    s->grid.cs = cs;
    s->grid.gs = s->prm.grid_size;
    XDEM_TRY(validate_params(s->prm));
    return grid_alloc(&s->grid, s->np + s->ntri);
}

// The geometry motion of step `s->step` (xDynamicsSolver.cpp:170-184: the
// time ct = step*dt is advanced, then every object whose motion has started
// is moved before oneStepAnalysis).  Moved planes and object poses are stored
// by one-lane kernels taking the struct by value; the triangles of every
// object are then re-posed from their local coordinates
// (updatePolygonObjectData_kernel, mphysics_cuda_impl.cuh:1911-1956).
static int32_t sim_apply_motions(xdem_sim* s, hipStream_t st) {
    const double dt = s->prm.dt;
    const double t = (double)s->step * dt;
    bool repose = false;
    for (Motion& mo : s->motions) {
        const xdem_motion& m = mo.m;
        bool active = t >= m.start_time && (!(m.end_time > m.start_time) || t <= m.end_time);
        if (active) mo.n_act++;
        const double tau = (double)mo.n_act * dt;
        if (mo.kind == XDEM_OBJECT_PLANE) {
            if (!active) continue;
            xdem_device_plane_info pl = s->h_planes[mo.index];
            auto mv = [&](xdem_double3& p) {
                p = xdem_double3{p.x + tau * m.velocity.x, p.y + tau * m.velocity.y,
                                 p.z + tau * m.velocity.z};
            };
            mv(pl.xw);
            mv(pl.w2);
            mv(pl.w3);
            mv(pl.w4);
            XDEM_TRY(launch_put_plane(s->d_planes + mo.index, pl, st));
        } else {
            if (!active && mo.active == active) continue;
            const xdem_device_polygon_mass_info& b = s->h_obj[mo.index];
            xdem_device_polygon_mass_info o = b;
            motion_pose(b.origin, b.ep, m, tau, o.origin, o.ep);
            o.vel = active ? m.velocity : xdem_double3{0, 0, 0};
            o.omega = active ? m.omega : xdem_double3{0, 0, 0};
            o.force = xdem_double3{0, 0, 0};
            o.moment = xdem_double3{0, 0, 0};
            XDEM_TRY(launch_put_object(s->d_obj + mo.index, o, st));
            repose = true;
        }
        mo.active = active;
    }
    if (repose) XDEM_TRY(launch_update_polygon(s->d_obj, s->d_vlist, s->d_sph, s->d_tri, s->ntri, st));
    return XDEM_OK;
}

// One dem_simulation::oneStepAnalysis (dem_simulation.cpp:447-469) on the GPU.
static int32_t sim_step_once(xdem_sim* s, hipStream_t st) {
    const KParams k = s->kparams();
    const int tcm = s->desc.tcm;
    PhaseTimer& T = s->timer;
    s->step++;
    // dem_simulation::oneStepAnalysis (dem_simulation.cpp:449-450): at step k with
    // (k - 1) % per_np == 0 the next group joins the simulated prefix (the group
    // list is walked as particleManager::NextCreatingPerGroup does,
    // particleManager.cpp:132-139: 0 once it is exhausted)
    if (s->release_every && (s->step - 1) % s->release_every == 0 && s->np_act < s->np)
        s->np_act = std::min(s->np, s->np_act + s->next_group());
    const uint32_t np = s->active();
    if (!s->motions.empty()) XDEM_TRY(sim_apply_motions(s, st));
    XDEM_TRY(T.begin(0, st));
    XDEM_TRY(launch_vv_position(k, s->pos, s->vel, s->acc, np, st));
    XDEM_TRY(T.end(st));
    XDEM_TRY(grid_detect(&s->grid, s->pos, s->ntri ? s->d_sph : nullptr, np, s->ntri, st, &T));
    const uint32_t* sid = s->grid.d_sorted_id;
    const uint32_t* cs0 = s->grid.d_cell_start;
    const uint32_t* ce0 = s->grid.d_cell_end;
    // the sorted screening copy (grown with the triangle spheres)
    const uint32_t nsorted = np + s->ntri;
    if (nsorted > s->pfs_cap) {
        if (s->d_pfs) (void)hipFree(s->d_pfs);
        s->d_pfs = nullptr;
        s->pfs_cap = 0;
        XDEM_HIP(hipMalloc(&s->d_pfs, sizeof(float4) * nsorted));
        s->pfs_cap = nsorted;
    }
    XDEM_TRY(T.begin(4, st));
    XDEM_TRY(launch_p2p(tcm, k, s->pos, s->vel, s->omega, s->force, s->moment, s->mass, sid, cs0,
                        ce0, s->d_cp_pp, np, nsorted, st, s->d_pfs));
    XDEM_TRY(T.end(st));
    XDEM_TRY(T.begin(5, st));
    for (const WallOp& op : s->ops) {
        if (op.kind == 0)
            XDEM_TRY(launch_planes(tcm, k, s->d_planes + op.first, s->d_plane_cp + op.first,
                                   op.count, s->pos, s->vel, s->omega, s->force, s->moment,
                                   s->mass, np, st));
        else if (op.kind == 1)
            XDEM_TRY(launch_cylinder(tcm, k, s->d_cyl + op.first, s->pos, s->vel, s->omega,
                                     s->force, s->moment, s->mass, np, s->d_cyl_cp + op.first,
                                     nullptr, nullptr, nullptr, st));
        else
            XDEM_TRY(launch_drum(tcm, k, s->d_drum + op.first, s->pos, s->vel, s->omega, s->force,
                                 s->moment, s->mass, np, s->d_drum_cp + op.first, nullptr,
                                 nullptr, st));
    }
    XDEM_TRY(T.end(st));
    if (s->ntri) {
        XDEM_TRY(T.begin(6, st));
        XDEM_TRY(launch_polygon(tcm, k, s->d_tri, s->d_obj, s->nobj, s->pos, s->vel, s->omega,
                                s->force, s->moment, s->mass, sid, cs0, ce0, s->d_obj_cp, np, st));
        XDEM_TRY(T.end(st));
    }
    XDEM_TRY(T.begin(7, st));
    XDEM_TRY(launch_vv_velocity(k, s->vel, s->acc, s->omega, s->alpha, s->force, s->moment,
                                s->mass, s->iner, np, st));
    XDEM_TRY(T.end(st));
    return XDEM_OK;
}

extern "C" {

// dem_simulation::initialize (dem_simulation.cpp:204-325): mass/inertia from
// radius and density, F = m g, zero velocities and accelerations.
int32_t xdem_sim_create(const xdem_sim_desc* d, const double* host_pos4, xdem_sim** out) {
    XDEM_REQUIRE(d && out && (d->np == 0 || host_pos4), "null argument");
    XDEM_REQUIRE(d->tcm == 0 || d->tcm == 1, "tcm must be 0 or 1");
    XDEM_REQUIRE(d->dt > 0, "dt must be > 0");
    *out = nullptr;
    xdem_sim* s = new xdem_sim();
    s->desc = *d;
    uint32_t np = d->np;
    s->np = np;
    std::vector<double> mass(np), iner(np), force(3 * (size_t)np);
    for (uint32_t i = 0; i < np; ++i) {
        double r = host_pos4[4 * (size_t)i + 3];
        mass[i] = d->particle.density * (4.0 / 3.0) * M_PI * std::pow(r, 3.0);
        iner[i] = (2.0 / 5.0) * mass[i] * std::pow(r, 2.0);
        force[3 * i + 0] = mass[i] * d->gravity.x;
        force[3 * i + 1] = mass[i] * d->gravity.y;
        force[3 * i + 2] = mass[i] * d->gravity.z;
        s->max_radius = std::max(s->max_radius, r);
    }
    s->prm.rollingCondition = d->rolling;
    s->prm.np = np;
    s->prm.nsphere = 0;
    s->prm.grid_size = d->grid_size;
    s->prm.ncell = d->grid_size.x * d->grid_size.y * d->grid_size.z;
    s->prm.dt = d->dt;
    s->prm.half2dt = 0.5 * d->dt * d->dt;
    s->prm.cohesion = 0.0;
    s->prm.gravity = d->gravity;
    s->prm.world_origin = d->world_origin;
    int32_t rc = sim_regrid(s);
    if (rc) {
        delete s;
        return rc;
    }
    size_t n3 = 3 * sizeof(double) * std::max<uint32_t>(np, 1);
    auto al = [&](double** p, size_t b) -> int32_t {
        XDEM_HIP(hipMalloc(p, b));
        XDEM_HIP(hipMemset(*p, 0, b));
        return XDEM_OK;
    };
    if ((rc = al(&s->pos, 4 * sizeof(double) * std::max<uint32_t>(np, 1))) ||
        (rc = al(&s->vel, n3)) || (rc = al(&s->omega, n3)) || (rc = al(&s->acc, n3)) ||
        (rc = al(&s->alpha, n3)) || (rc = al(&s->force, n3)) || (rc = al(&s->moment, n3)) ||
        (rc = al(&s->mass, sizeof(double) * std::max<uint32_t>(np, 1))) ||
        (rc = al(&s->iner, sizeof(double) * std::max<uint32_t>(np, 1)))) {
        delete s;
        return rc;
    }
    auto h2d = [&](void* dst, const void* src, size_t b) -> int32_t {
        if (b) XDEM_HIP(hipMemcpy(dst, src, b, hipMemcpyHostToDevice));
        return XDEM_OK;
    };
    if ((rc = h2d(s->pos, host_pos4, 4 * sizeof(double) * np)) ||
        (rc = h2d(s->mass, mass.data(), sizeof(double) * np)) ||
        (rc = h2d(s->iner, iner.data(), sizeof(double) * np)) ||
        (rc = h2d(s->force, force.data(), 3 * sizeof(double) * np))) {
        delete s;
        return rc;
    }
    std::vector<xdem_device_contact_property> cp = {make_cp(d->particle, d->pp)};
    if ((rc = upload(&s->d_cp_pp, cp))) {
        delete s;
        return rc;
    }
    *out = s;
    return XDEM_OK;
}

int32_t xdem_sim_destroy(xdem_sim* s) {
    if (s) (void)hipStreamSynchronize(stream());
    delete s;
    return XDEM_OK;
}

static void push_op(xdem_sim* s, int kind, uint32_t idx) {
    if (kind == 0 && !s->ops.empty() && s->ops.back().kind == 0 &&
        s->ops.back().first + s->ops.back().count == idx)
        s->ops.back().count++;
    else
        s->ops.push_back(WallOp{kind, idx, 1});
}

int32_t xdem_sim_add_plane(xdem_sim* s, const xdem_device_plane_info* plane,
                           const xdem_contact_pair_desc* c) {
    XDEM_REQUIRE(s && plane && c, "null argument");
    s->h_planes.push_back(*plane);
    s->h_plane_cp.push_back(make_cp(s->desc.particle, *c));
    push_op(s, 0, (uint32_t)s->h_planes.size() - 1);
    s->dirty = true;
    return XDEM_OK;
}

int32_t xdem_sim_add_cube(xdem_sim* s, const xdem_device_plane_info* planes6,
                          const xdem_contact_pair_desc* c) {
    XDEM_REQUIRE(s && planes6 && c, "null argument");
    for (int i = 0; i < 6; ++i) XDEM_TRY(xdem_sim_add_plane(s, planes6 + i, c));
    return XDEM_OK;
}

int32_t xdem_sim_add_cylinder(xdem_sim* s, const xdem_device_cylinder_info* cyl,
                              const xdem_contact_pair_desc* c) {
    XDEM_REQUIRE(s && cyl && c, "null argument");
    s->h_cyl.push_back(*cyl);
    s->h_cyl_cp.push_back(make_cp(s->desc.particle, *c));
    push_op(s, 1, (uint32_t)s->h_cyl.size() - 1);
    s->dirty = true;
    return XDEM_OK;
}

int32_t xdem_sim_add_drum(xdem_sim* s, const xdem_drum_info* drum, const xdem_contact_pair_desc* c) {
    XDEM_REQUIRE(s && drum && c, "null argument");
    const xdem_double3& a = drum->axis;
    double an = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    XDEM_REQUIRE(std::fabs(an - 1.0) < 1e-12, "drum axis must be a unit vector (|axis| = %.17g)", an);
    XDEM_REQUIRE(drum->radius > 0 && drum->half_length > 0, "drum radius and half length must be > 0");
    s->h_drum.push_back(*drum);
    s->h_drum_cp.push_back(make_cp(s->desc.particle, *c));
    push_op(s, 2, (uint32_t)s->h_drum.size() - 1);
    s->dirty = true;
    return XDEM_OK;
}

// contact_particles_polygonObjects::define (contact_particles_polygonObjects.cpp:47-123):
// triangles with their circumspheres are appended to the hash arrays at sid = np.
int32_t xdem_sim_set_polygons(xdem_sim* s, const xdem_device_polygon_info* tris,
                              const double* sphere4, uint32_t ntri,
                              const xdem_device_polygon_mass_info* objs,
                              const xdem_contact_pair_desc* per_object, uint32_t nobj) {
    XDEM_REQUIRE(s && (ntri == 0 || (tris && sphere4 && objs && per_object)), "null argument");
    XDEM_REQUIRE(ntri == 0 || nobj > 0, "triangles need at least one object");
    for (uint32_t t = 0; t < ntri; ++t)
        XDEM_REQUIRE(tris[t].id >= 0 && (uint32_t)tris[t].id < nobj,
                     "triangle %u has object id %d outside [0,%u)", t, tris[t].id, nobj);
    std::vector<xdem_device_polygon_info> ht(tris, tris + ntri);
    std::vector<double> hs(sphere4, sphere4 + 4 * (size_t)ntri);
    std::vector<xdem_device_polygon_mass_info> ho(objs, objs + nobj);
    std::vector<xdem_device_contact_property> hc;
    for (uint32_t o = 0; o < nobj; ++o) hc.push_back(make_cp(s->desc.particle, per_object[o]));
    XDEM_TRY(upload(&s->d_tri, ht));
    XDEM_TRY(upload(&s->d_sph, hs));
    XDEM_TRY(upload(&s->d_obj, ho));
    XDEM_TRY(upload(&s->d_obj_cp, hc));
    // local vertex coordinates, A(ep0)^T (P - origin0) as toLocal
    // (mphysics_cuda_impl.cuh:640-645), for objects that are later moved
    s->h_vlist.assign(9 * (size_t)ntri, 0.0);
    for (uint32_t t = 0; t < ntri; ++t) {
        const xdem_device_polygon_mass_info& o = ho[tris[t].id];
        d3 r0, r1, r2;
        ep_rows(o.ep, r0, r1, r2);
        const xdem_double3* v[3] = {&tris[t].P, &tris[t].Q, &tris[t].R};
        for (int j = 0; j < 3; ++j) {
            d3 d = mk3(v[j]->x - o.origin.x, v[j]->y - o.origin.y, v[j]->z - o.origin.z);
            double* l = &s->h_vlist[9 * (size_t)t + 3 * j];
            l[0] = r0.x * d.x + r1.x * d.y + r2.x * d.z;
            l[1] = r0.y * d.x + r1.y * d.y + r2.y * d.z;
            l[2] = r0.z * d.x + r1.z * d.y + r2.z * d.z;
        }
    }
    XDEM_TRY(upload(&s->d_vlist, s->h_vlist));
    s->h_obj = ho;
    s->motions.erase(std::remove_if(s->motions.begin(), s->motions.end(),
                                    [](const Motion& m) { return m.kind == XDEM_OBJECT_POLYGON; }),
                     s->motions.end());
    s->ntri = ntri;
    s->nobj = nobj;
    // dem_simulation.cpp:226-227,249-250: the cell size covers the largest circumsphere.
    for (uint32_t t = 0; t < ntri; ++t) s->max_radius = std::max(s->max_radius, hs[4 * t + 3]);
    return sim_regrid(s);
}

int32_t xdem_sim_set_motion(xdem_sim* s, int32_t kind, uint32_t index, const xdem_motion* m) {
    XDEM_REQUIRE(s && m, "null argument");
    XDEM_REQUIRE(kind == XDEM_OBJECT_PLANE || kind == XDEM_OBJECT_POLYGON,
                 "kind must be XDEM_OBJECT_PLANE or XDEM_OBJECT_POLYGON");
    if (kind == XDEM_OBJECT_PLANE) {
        XDEM_REQUIRE(index < s->h_planes.size(), "plane %u out of range [0,%zu)", index,
                     s->h_planes.size());
        XDEM_REQUIRE(m->omega.x == 0 && m->omega.y == 0 && m->omega.z == 0,
                     "a plane moves by translation only");
    } else {
        XDEM_REQUIRE(index < s->nobj, "polygon object %u out of range [0,%u)", index, s->nobj);
    }
    for (Motion& mo : s->motions)
        if (mo.kind == kind && mo.index == index) {
            mo.m = *m;
            return XDEM_OK;
        }
    s->motions.push_back(Motion{kind, index, *m, 0, true});
    return XDEM_OK;
}

int32_t xdem_sim_get_polygons(const xdem_sim* s, xdem_device_polygon_info* tris, double* sphere4) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_HIP(hipStreamSynchronize(stream()));
    if (tris && s->ntri)
        XDEM_HIP(hipMemcpy(tris, s->d_tri, sizeof(xdem_device_polygon_info) * s->ntri,
                           hipMemcpyDeviceToHost));
    if (sphere4 && s->ntri)
        XDEM_HIP(hipMemcpy(sphere4, s->d_sph, 4 * sizeof(double) * s->ntri, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int32_t xdem_sim_get_objects(const xdem_sim* s, xdem_device_polygon_mass_info* objs) {
    XDEM_REQUIRE(s && (objs || !s->nobj), "null argument");
    XDEM_HIP(hipStreamSynchronize(stream()));
    if (s->nobj)
        XDEM_HIP(hipMemcpy(objs, s->d_obj, sizeof(xdem_device_polygon_mass_info) * s->nobj,
                           hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int32_t xdem_sim_set_object_state(xdem_sim* s, uint32_t obj,
                                  const xdem_device_polygon_mass_info* state) {
    XDEM_REQUIRE(s && state, "null argument");
    XDEM_REQUIRE(obj < s->nobj, "polygon object %u out of range [0,%u)", obj, s->nobj);
    for (const Motion& m : s->motions)
        XDEM_REQUIRE(!(m.kind == XDEM_OBJECT_POLYGON && m.index == obj),
                     "object %u has a prescribed motion", obj);
    xdem_device_polygon_mass_info o = *state;
    o.force = xdem_double3{0, 0, 0};
    o.moment = xdem_double3{0, 0, 0};
    hipStream_t st = stream();
    XDEM_TRY(launch_put_object(s->d_obj + obj, o, st));
    XDEM_TRY(launch_update_polygon(s->d_obj, s->d_vlist, s->d_sph, s->d_tri, s->ntri, st));
    return XDEM_OK;
}

uint64_t xdem_sim_steps(const xdem_sim* s) { return s ? s->step : 0; }

int32_t xdem_sim_step(xdem_sim* s, uint32_t nsteps) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_TRY(sim_sync_walls(s));
    hipStream_t st = stream();
    for (uint32_t i = 0; i < nsteps; ++i) XDEM_TRY(sim_step_once(s, st));
    return XDEM_OK;
}

int32_t xdem_sim_get_state(const xdem_sim* s, double* pos4, double* vel3, double* omega3,
                           double* acc3, double* alpha3, double* force3, double* moment3) {
    XDEM_REQUIRE(s, "null sim");
    hipStream_t st = stream();
    XDEM_HIP(hipStreamSynchronize(st));
    size_t n = s->np;
    auto cp = [&](double* dst, const double* src, size_t b) -> int32_t {
        if (dst && b) XDEM_HIP(hipMemcpy(dst, src, b, hipMemcpyDeviceToHost));
        return XDEM_OK;
    };
    XDEM_TRY(cp(pos4, s->pos, 32 * n));
    XDEM_TRY(cp(vel3, s->vel, 24 * n));
    XDEM_TRY(cp(omega3, s->omega, 24 * n));
    XDEM_TRY(cp(acc3, s->acc, 24 * n));
    XDEM_TRY(cp(alpha3, s->alpha, 24 * n));
    XDEM_TRY(cp(force3, s->force, 24 * n));
    XDEM_TRY(cp(moment3, s->moment, 24 * n));
    return XDEM_OK;
}

int32_t xdem_sim_set_state(xdem_sim* s, const double* pos4, const double* vel3,
                           const double* omega3, const double* acc3, const double* alpha3,
                           const double* force3, const double* moment3) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_HIP(hipStreamSynchronize(stream()));
    size_t n = s->np;
    auto cp = [&](double* dst, const double* src, size_t b) -> int32_t {
        if (src && b) XDEM_HIP(hipMemcpy(dst, src, b, hipMemcpyHostToDevice));
        return XDEM_OK;
    };
    XDEM_TRY(cp(s->pos, pos4, 32 * n));
    XDEM_TRY(cp(s->vel, vel3, 24 * n));
    XDEM_TRY(cp(s->omega, omega3, 24 * n));
    XDEM_TRY(cp(s->acc, acc3, 24 * n));
    XDEM_TRY(cp(s->alpha, alpha3, 24 * n));
    XDEM_TRY(cp(s->force, force3, 24 * n));
    XDEM_TRY(cp(s->moment, moment3, 24 * n));
    return XDEM_OK;
}

int32_t xdem_sim_device_arrays(xdem_sim* s, double** pos, double** vel, double** omega,
                               double** acc, double** alpha, double** force, double** moment,
                               double** mass, double** iner) {
    XDEM_REQUIRE(s, "null sim");
    if (pos) *pos = s->pos;
    if (vel) *vel = s->vel;
    if (omega) *omega = s->omega;
    if (acc) *acc = s->acc;
    if (alpha) *alpha = s->alpha;
    if (force) *force = s->force;
    if (moment) *moment = s->moment;
    if (mass) *mass = s->mass;
    if (iner) *iner = s->iner;
    return XDEM_OK;
}

int32_t xdem_sim_enable_timing(xdem_sim* s, int32_t on) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_TRY(s->timer.flush());
    s->timer.reset();
    s->timer.on = on != 0;
    return XDEM_OK;
}

int32_t xdem_sim_phase_times(xdem_sim* s, double* ms8, uint32_t* count8) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_TRY(s->timer.flush());
    for (int i = 0; i < PhaseTimer::NPHASE; ++i) {
        if (ms8) ms8[i] = s->timer.ms[i];
        if (count8) count8[i] = s->timer.count[i];
    }
    return XDEM_OK;
}

uint32_t xdem_sim_np(const xdem_sim* s) { return s ? s->np : 0; }

uint32_t xdem_sim_active(const xdem_sim* s) { return s ? s->active() : 0; }

int32_t xdem_sim_set_release(xdem_sim* s, uint32_t every_steps, uint32_t group) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_REQUIRE(s->step == 0, "the release is set before the first step");
    XDEM_REQUIRE(every_steps == 0 || group > 0, "a release needs a group size > 0");
    s->release_every = every_steps;
    s->release_group = every_steps ? group : 0;
    s->release_groups.clear();
    s->release_next = 0;
    s->np_act = 0;
    return XDEM_OK;
}

int32_t xdem_sim_set_release_groups(xdem_sim* s, uint32_t every_steps, const uint32_t* groups,
                                    uint32_t ngroups) {
    XDEM_REQUIRE(s, "null sim");
    XDEM_REQUIRE(s->step == 0, "the release is set before the first step");
    XDEM_REQUIRE(every_steps > 0 && groups && ngroups > 0,
                 "a group release needs every_steps > 0 and a non-empty group list");
    uint64_t total = 0;
    for (uint32_t g = 0; g < ngroups; ++g) total += groups[g];
    XDEM_REQUIRE(total <= s->np, "the groups hold %llu particles, the simulation %u",
                 (unsigned long long)total, s->np);
    s->release_every = every_steps;
    s->release_group = 0;
    s->release_groups.assign(groups, groups + ngroups);
    s->release_next = 0;
    s->np_act = 0;
    return XDEM_OK;
}
xdem_grid* xdem_sim_grid(xdem_sim* s) { return s ? &s->grid : nullptr; }

// ------------------------------------------------------------ state files
int32_t xdem_sim_save_part(const xdem_sim* s, const char* dir, uint32_t part) {
    XDEM_REQUIRE(s && dir, "null argument");
    char path[4096];
    int len = std::snprintf(path, sizeof(path), "%s/part%04u.bin", dir, part);
    XDEM_REQUIRE(len > 0 && len < (int)sizeof(path), "path too long");
    std::vector<double> pos(4 * (size_t)s->np);
    XDEM_TRY(xdem_sim_get_state(s, pos.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    return xdem_write_part_f64(path, (double)s->step * s->prm.dt, s->np, pos.data());
}

int32_t xdem_sim_save_final(const xdem_sim* s, const char* path) {
    XDEM_REQUIRE(s, "null sim");
    size_t n = s->np;
    std::vector<double> pos(4 * n), vel(3 * n), om(3 * n);
    XDEM_TRY(xdem_sim_get_state(s, pos.data(), vel.data(), om.data(), nullptr, nullptr, nullptr,
                                nullptr));
    return xdem_write_bfr_f64(path, (double)s->step * s->prm.dt, s->np, pos.data(), vel.data(),
                              om.data());
}

// dem_simulation::initialize (force = m g, zero accelerations) followed by
// setStartingData (positions, velocities, angular velocities).
int32_t xdem_sim_load_bfr(xdem_sim* s, const char* path) {
    XDEM_REQUIRE(s, "null sim");
    size_t n = s->np;
    double ct = 0;
    int32_t prec = 0;
    std::vector<double> pos(4 * n), vel(3 * n), om(3 * n), zero(3 * n, 0.0), mass(n), f(3 * n);
    XDEM_TRY(xdem_read_bfr(path, s->np, &ct, &prec, pos.data(), vel.data(), om.data()));
    XDEM_HIP(hipStreamSynchronize(stream()));
    if (n) XDEM_HIP(hipMemcpy(mass.data(), s->mass, 8 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        f[3 * i + 0] = mass[i] * s->desc.gravity.x;
        f[3 * i + 1] = mass[i] * s->desc.gravity.y;
        f[3 * i + 2] = mass[i] * s->desc.gravity.z;
    }
    XDEM_TRY(xdem_sim_set_state(s, pos.data(), vel.data(), om.data(), zero.data(), zero.data(),
                                f.data(), zero.data()));
    s->step = (uint64_t)std::llround(ct / s->prm.dt);
    return XDEM_OK;
}

// Payload: np, ntri, nobj, nplane, nmotion, step, dt; the seven particle
// arrays; the current planes, triangles, circumspheres and objects; per
// motion its kind, index, active-step count, active flag and parameters.
int32_t xdem_sim_save_checkpoint(const xdem_sim* s, const char* path) {
    XDEM_REQUIRE(s, "null sim");
    size_t n = s->np;
    std::vector<double> a[7];
    for (int k = 0; k < 7; ++k) a[k].resize((k == 0 ? 4 : 3) * n);
    XDEM_TRY(xdem_sim_get_state(s, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                                a[5].data(), a[6].data()));
    const uint32_t nplane = s->dirty ? 0u : (uint32_t)s->h_planes.size();
    std::vector<xdem_device_plane_info> planes(nplane);
    std::vector<xdem_device_polygon_info> tris(s->ntri);
    std::vector<double> sph(4 * (size_t)s->ntri);
    std::vector<xdem_device_polygon_mass_info> objs(s->nobj);
    if (nplane)
        XDEM_HIP(hipMemcpy(planes.data(), s->d_planes, sizeof(xdem_device_plane_info) * nplane,
                           hipMemcpyDeviceToHost));
    XDEM_TRY(xdem_sim_get_polygons(s, tris.data(), sph.data()));
    XDEM_TRY(xdem_sim_get_objects(s, objs.data()));
    io::Writer w;
    w.put(s->np);
    w.put(s->ntri);
    w.put(s->nobj);
    w.put(nplane);
    w.put((uint32_t)s->motions.size());
    w.put(s->step);
    w.put(s->prm.dt);
    w.put(s->release_every);
    w.put(s->release_group);
    w.put(s->np_act);
    w.put(s->release_next);
    w.put((uint32_t)s->release_groups.size());
    w.arr(s->release_groups.data(), s->release_groups.size());
    for (int k = 0; k < 7; ++k) w.arr(a[k].data(), a[k].size());
    w.arr(planes.data(), planes.size());
    w.arr(tris.data(), tris.size());
    w.arr(sph.data(), sph.size());
    w.arr(objs.data(), objs.size());
    for (const Motion& m : s->motions) {
        w.put(m.kind);
        w.put(m.index);
        w.put(m.n_act);
        w.put((int32_t)(m.active ? 1 : 0));
        w.put(m.m);
    }
    return io::write_checkpoint(path, io::CKPT_SIM, w);
}

int32_t xdem_sim_load_checkpoint(xdem_sim* s, const char* path) {
    XDEM_REQUIRE(s, "null sim");
    std::vector<uint8_t> file;
    io::Reader r;
    XDEM_TRY(io::read_checkpoint(path, io::CKPT_SIM, file, r, nullptr));
    uint32_t np, ntri, nobj, nplane, nmot, rel_every, rel_group, np_act, rel_next, ngroups;
    uint64_t step;
    double dt;
    XDEM_REQUIRE(r.get(np) && r.get(ntri) && r.get(nobj) && r.get(nplane) && r.get(nmot) &&
                     r.get(step) && r.get(dt) && r.get(rel_every) && r.get(rel_group) &&
                     r.get(np_act) && r.get(rel_next) && r.get(ngroups),
                 "%s: short header", path);
    XDEM_REQUIRE(np_act <= np, "%s: %u active of %u particles", path, np_act, np);
    XDEM_REQUIRE(ngroups <= np && rel_next <= ngroups, "%s: release group %u of %u", path,
                 rel_next, ngroups);
    std::vector<uint32_t> groups(ngroups);
    XDEM_REQUIRE(r.arr(groups.data(), ngroups), "%s: truncated release groups", path);
    XDEM_REQUIRE(np == s->np && ntri == s->ntri && nobj == s->nobj &&
                     nmot == (uint32_t)s->motions.size(),
                 "%s: built for %u particles, %u triangles, %u objects, %u motions "
                 "(this simulation: %u, %u, %u, %zu)",
                 path, np, ntri, nobj, nmot, s->np, s->ntri, s->nobj, s->motions.size());
    XDEM_REQUIRE(nplane == 0 || nplane == (uint32_t)s->h_planes.size(), "%s: %u planes, not %zu",
                 path, nplane, s->h_planes.size());
    XDEM_REQUIRE(dt == s->prm.dt, "%s: dt %.17g differs from %.17g", path, dt, s->prm.dt);
    size_t n = np;
    std::vector<double> a[7];
    for (int k = 0; k < 7; ++k) {
        a[k].resize((k == 0 ? 4 : 3) * n);
        XDEM_REQUIRE(r.arr(a[k].data(), a[k].size()), "%s: truncated state", path);
    }
    std::vector<xdem_device_plane_info> planes(nplane);
    std::vector<xdem_device_polygon_info> tris(ntri);
    std::vector<double> sph(4 * (size_t)ntri);
    std::vector<xdem_device_polygon_mass_info> objs(nobj);
    XDEM_REQUIRE(r.arr(planes.data(), nplane) && r.arr(tris.data(), ntri) &&
                     r.arr(sph.data(), sph.size()) && r.arr(objs.data(), nobj),
                 "%s: truncated geometry", path);
    std::vector<Motion> mot(s->motions);
    for (Motion& m : mot) {
        int32_t kind, active;
        uint32_t index;
        XDEM_REQUIRE(r.get(kind) && r.get(index) && r.get(m.n_act) && r.get(active) && r.get(m.m),
                     "%s: truncated motions", path);
        XDEM_REQUIRE(kind == m.kind && index == m.index, "%s: motion (%d, %u) is (%d, %u) here",
                     path, kind, index, m.kind, m.index);
        m.active = active != 0;
    }
    XDEM_REQUIRE(r.done(), "%s: trailing bytes", path);
    XDEM_TRY(sim_sync_walls(s));
    XDEM_TRY(xdem_sim_set_state(s, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                a[4].data(), a[5].data(), a[6].data()));
    if (nplane)
        XDEM_HIP(hipMemcpy(s->d_planes, planes.data(), sizeof(xdem_device_plane_info) * nplane,
                           hipMemcpyHostToDevice));
    if (ntri) {
        XDEM_HIP(hipMemcpy(s->d_tri, tris.data(), sizeof(xdem_device_polygon_info) * ntri,
                           hipMemcpyHostToDevice));
        XDEM_HIP(hipMemcpy(s->d_sph, sph.data(), sizeof(double) * sph.size(),
                           hipMemcpyHostToDevice));
    }
    if (nobj)
        XDEM_HIP(hipMemcpy(s->d_obj, objs.data(), sizeof(xdem_device_polygon_mass_info) * nobj,
                           hipMemcpyHostToDevice));
    s->motions = mot;
    s->step = step;
    s->release_every = rel_every;
    s->release_group = rel_group;
    s->np_act = np_act;
    s->release_next = rel_next;
    s->release_groups = std::move(groups);
    return XDEM_OK;
}

}  // extern "C"
