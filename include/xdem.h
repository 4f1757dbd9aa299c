/*
 * xdem.h — C ABI of the MI355X (gfx950) DEM contact hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int32_t status (XDEM_OK == 0).  xdem_last_error() returns the message of
 * the last failure on the calling thread.
 *
 * Device pointers are caller-owned, as in the reference: the library never frees
 * caller memory.  xdem_malloc/xdem_free/xdem_memcpy_* are provided so that a
 * caller without a HIP toolchain can manage buffers.
 *
 * Struct layouts reproduce the reference's CUDA vector_types alignment
 * (double3: 24 B / 8-aligned, double4: 32 B / 16-aligned) and are checked with
 * static_assert below.
 *
 * Reference interfaces replaced (citations relative to the reference tree):
 *   Set A (live)  include/mphysics_cuda_dec.cuh:17-161 (FP64)
 *   Set B (older) include/cu_dem_dec.cuh:41-113
 *   grid_base     include/grid_base.h:11-46
 *   integrator    include/dem_integrator.h:16-28
 */
#ifndef XDEM_H
#define XDEM_H

#include <stdint.h>
#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    XDEM_OK = 0,
    XDEM_ERR_INVALID = 1,     /* bad argument (null pointer, size, grid not power of two ...) */
    XDEM_ERR_HIP = 2,         /* a HIP runtime call failed */
    XDEM_ERR_NO_DEVICE = 3,   /* no gfx950 device visible */
    XDEM_ERR_CAPACITY = 4,    /* request exceeds a fixed capacity (batch slots, HBM budget) */
    XDEM_ERR_STATE = 5        /* call out of order (e.g. parameters never set) */
};

/* ------------------------------------------------------- reference POD types */
typedef struct xdem_uint3 { uint32_t x, y, z; } xdem_uint3;
typedef struct xdem_double3 { double x, y, z; } xdem_double3;
typedef struct xdem_double4 { double x, y, z, w; } __attribute__((aligned(16))) xdem_double4;

/* device_parameters — include/mphysics_cuda_dec.cuh:17-30 */
typedef struct xdem_device_parameters {
    bool rollingCondition;
    uint32_t np;
    uint32_t nsphere;
    uint32_t ncell;
    xdem_uint3 grid_size;
    double dt;
    double half2dt;
    double cell_size;
    double cohesion;
    xdem_double3 gravity;
    xdem_double3 world_origin;
} xdem_device_parameters;

/* device_polygon_info — include/mphysics_cuda_dec.cuh:32-41 */
typedef struct xdem_device_polygon_info {
    int32_t id;
    xdem_double3 P, Q, R, V, W, N;
} xdem_device_polygon_info;

/* device_plane_info — include/mphysics_cuda_dec.cuh:43-55 */
typedef struct xdem_device_plane_info {
    double l1, l2;
    xdem_double3 u1, u2, uw, xw, pa, pb, w2, w3, w4;
} xdem_device_plane_info;

/* device_polygon_mass_info — include/mphysics_cuda_dec.cuh:57-65 */
typedef struct xdem_device_polygon_mass_info {
    xdem_double3 origin;
    xdem_double4 ep;
    xdem_double3 vel, omega, force, moment;
} xdem_device_polygon_mass_info;

/* device_cylinder_info — include/mphysics_cuda_dec.cuh:67-76 */
typedef struct xdem_device_cylinder_info {
    double len, rbase, rtop;
    xdem_double3 pbase, ptop, origin, vel, omega;
    xdem_double4 ep;
} xdem_device_cylinder_info;

/* device_contact_property — include/mphysics_cuda_dec.cuh:78-88 */
typedef struct xdem_device_contact_property {
    double Ei, Ej, pri, prj, Gi, Gj, rest, fric, rfric, coh, sratio;
} xdem_device_contact_property;

/* device_force_constant — include/mphysics_cuda_dec.cuh:91-99 */
typedef struct xdem_device_force_constant {
    double kn, vn, ks, vs, mu, ms;
} xdem_device_force_constant;

#ifdef __cplusplus
static_assert(sizeof(xdem_device_parameters) == 112, "device_parameters layout");
static_assert(offsetof(xdem_device_parameters, dt) == 32, "device_parameters.dt");
static_assert(offsetof(xdem_device_parameters, world_origin) == 88, "device_parameters.world_origin");
static_assert(sizeof(xdem_device_polygon_info) == 152, "device_polygon_info layout");
static_assert(sizeof(xdem_device_plane_info) == 232, "device_plane_info layout");
static_assert(sizeof(xdem_device_polygon_mass_info) == 160, "device_polygon_mass_info layout");
static_assert(offsetof(xdem_device_polygon_mass_info, ep) == 32, "device_polygon_mass_info.ep");
static_assert(sizeof(xdem_device_cylinder_info) == 176, "device_cylinder_info layout");
static_assert(offsetof(xdem_device_cylinder_info, ep) == 144, "device_cylinder_info.ep");
static_assert(sizeof(xdem_device_contact_property) == 88, "device_contact_property layout");
#endif

/* ------------------------------------------------------------ housekeeping */
const char* xdem_last_error(void);
int32_t xdem_version(void);
/* Number of visible gfx950 devices (0 without a GPU; never fails). */
int32_t xdem_device_count(void);
int32_t xdem_synchronize(void);
int32_t xdem_malloc(void** dptr, size_t bytes);
int32_t xdem_free(void* dptr);
int32_t xdem_memset(void* dptr, int value, size_t bytes);
int32_t xdem_memcpy_htod(void* dst, const void* src, size_t bytes);
int32_t xdem_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int32_t xdem_memcpy_dtod(void* dst, const void* src, size_t bytes);
/* Free and total device memory of the current device (hipMemGetInfo). */
int32_t xdem_device_memory(size_t* free_bytes, size_t* total_bytes);
/* A plain copy kernel of `bytes` (multiple of 16, 16-byte aligned pointers):
 * the known-size reference for calibrating profiler byte counters.  No
 * counterpart in the reference. */
int32_t xdem_copy_kernel(void* dst, const void* src, size_t bytes);

/* ====================================================================
 * Set A — live FP64 API (include/mphysics_cuda_dec.cuh:110-161).
 * All pointers are device pointers.  Layout contract (reference):
 *   pos    double4{x,y,z,r} x np
 *   vel/acc/omega/alpha/force/moment  packed double3 x np
 *   mass/iner  double x np
 *   hash/index/sorted_index  uint32 x n (n = np + nsphere; particles first)
 *   cstart/cend  uint32 x ncell
 * Launches go to the library stream (the legacy null stream unless
 * xdem_set_stream is called), in call order, like the reference.
 * ==================================================================== */

/* setSymbolicParameter — mphysics_cuda_dec.cuh:110, def mphysics_cuda_def.cu:13-16.
 * grid_size components must be powers of two (the hash wraps by mask). */
int32_t xdem_set_symbolic_parameter(const xdem_device_parameters* h_paras);
int32_t xdem_get_symbolic_parameter(xdem_device_parameters* out);
/* Stream used by every launch of this library (hipStream_t; NULL = null stream). */
int32_t xdem_set_stream(void* hip_stream);

/* vv_update_position — mphysics_cuda_dec.cuh:112, def mphysics_cuda_def.cu:33-41 */
int32_t xdem_vv_update_position_f64(double* pos, double* vel, double* acc, uint32_t np);
/* vv_update_velocity — mphysics_cuda_dec.cuh:113, def mphysics_cuda_def.cu:43-56 */
int32_t xdem_vv_update_velocity_f64(double* vel, double* acc, double* omega, double* alpha,
                                    double* force, double* moment, double* mass, double* iner,
                                    uint32_t np);
/* cu_calculateHashAndIndex — mphysics_cuda_dec.cuh:115, def mphysics_cuda_def.cu:58-66 */
int32_t xdem_calculate_hash_and_index_f64(uint32_t* hash, uint32_t* index, const double* pos,
                                          uint32_t np);
/* cu_calculateHashAndIndexForPolygonSphere — mphysics_cuda_dec.cuh:116-118, def :68-77 */
int32_t xdem_calculate_hash_and_index_polygon_sphere_f64(uint32_t* hash, uint32_t* index,
                                                         uint32_t sid, uint32_t nsphere,
                                                         const double* sphere);
/* cu_reorderDataAndFindCellStart — mphysics_cuda_dec.cuh:119, def mphysics_cuda_def.cu:80-108.
 * Stable sort of (hash, index) by hash, then cstart/cend over all ncell cells
 * (empty cells: 0xFFFFFFFF / 0) and sorted_index.  hash and index are sorted in place. */
int32_t xdem_reorder_data_and_find_cell_start(uint32_t* hash, uint32_t* index, uint32_t* cstart,
                                              uint32_t* cend, uint32_t* sorted_index, uint32_t n,
                                              uint32_t ncell);
/* No counterpart in the reference: xdem_reorder_data_and_find_cell_start with
 * the same arguments and results, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It calls no allocating routine: its sort scratch is a
 * fixed device array of the library (about 8.7 MB, held by each device the
 * library runs on until the process ends), so n is limited to
 * XDEM_NOALLOC_MAX_KEYS (larger n: an error, nothing written).  All calls on
 * one device share that scratch: calls that overlap on one device (several
 * host threads, or streams other than the library's) race.  Every scratch
 * word is written before it is read in a call, so nothing carries over from
 * the previous call. */
#define XDEM_NOALLOC_MAX_KEYS (1u << 20)
int32_t xdem_reorder_data_and_find_cell_start_noalloc(uint32_t* hash, uint32_t* index,
                                                      uint32_t* cstart, uint32_t* cend,
                                                      uint32_t* sorted_index, uint32_t n,
                                                      uint32_t ncell);
/* cu_calculate_p2p — mphysics_cuda_dec.cuh:121-126, def mphysics_cuda_def.cu:110-136 */
int32_t xdem_calculate_p2p_f64(int32_t tcm, const double* pos, const double* vel,
                               const double* omega, double* force, double* moment,
                               const double* mass, const uint32_t* sorted_index,
                               const uint32_t* cstart, const uint32_t* cend,
                               const xdem_device_contact_property* cp, uint32_t np);
/* cu_plane_contact_force — mphysics_cuda_dec.cuh:129-133, def mphysics_cuda_def.cu:138-156 */
int32_t xdem_plane_contact_force_f64(int32_t tcm, const xdem_device_plane_info* plane,
                                     const double* pos, const double* vel, const double* omega,
                                     double* force, double* moment, const double* mass,
                                     uint32_t np, const xdem_device_contact_property* cp);
/* cu_cube_contact_force — mphysics_cuda_dec.cuh:135-139, def mphysics_cuda_def.cu:158-179
 * (six planes, applied in array order). */
int32_t xdem_cube_contact_force_f64(int32_t tcm, const xdem_device_plane_info* planes6,
                                    const double* pos, const double* vel, const double* omega,
                                    double* force, double* moment, const double* mass,
                                    uint32_t np, const xdem_device_contact_property* cp);
/* cu_particle_polygonObject_collision — mphysics_cuda_dec.cuh:142-147, def :203-221.
 * cp and dpmi are indexed by polygon-object id (dpi[t].id). The per-object
 * reaction sums dpmi[].force/moment are written by an ordered reduction
 * (no racing +=), so they are deterministic: dpmi[o].force = -sum (Fn + Ft)
 * and dpmi[o].moment = -sum (cp - origin) x (Fn + Ft) over the contacts of
 * this call, as the reference caller reads them back
 * (source/contact_particles_polygonObjects.cpp:469-477).  nobject = 0 leaves
 * dpmi untouched; nobject = XDEM_NOBJECT_AUTO takes the number of objects as
 * 1 + the largest dpi[t].id over the ntriangle triangles (ntriangle = 0: the
 * parameter block's nsphere, the triangle count the reference's caller
 * sets), which costs one device-to-host read. */
#define XDEM_NOBJECT_AUTO 0xFFFFFFFFu
int32_t xdem_particle_polygon_object_collision_f64(
    int32_t tcm, const xdem_device_polygon_info* dpi, const double* dsph,
    xdem_device_polygon_mass_info* dpmi, uint32_t nobject, uint32_t ntriangle,
    const double* pos, const double* vel, const double* omega, double* force, double* moment,
    const double* mass, const uint32_t* sorted_index, const uint32_t* cstart,
    const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np);
/* cu_cylinder_hertzian_contact_force — mphysics_cuda_dec.cuh:150-155, def :181-201.
 * mf/mm (device, np entries) receive the per-particle reaction; out_mf/out_mm
 * (host) receive their sums (the reference leaves _mf/_mm unwritten). */
int32_t xdem_cylinder_hertzian_contact_force_f64(
    int32_t tcm, const xdem_device_cylinder_info* cyl, const double* pos, const double* vel,
    const double* omega, double* force, double* moment, const double* mass, uint32_t np,
    const xdem_device_contact_property* cp, const xdem_double3* mpos, xdem_double3* mf,
    xdem_double3* mm, xdem_double3* out_mf, xdem_double3* out_mm);
/* Rotating drum: the inside of a cylindrical shell of radius `radius` around
 * the axis through `origin` along the unit vector `axis`, between the end
 * caps at axial coordinates -half_length and +half_length, spinning at
 * `omega` rad/s about +axis (right-hand rule).  No counterpart in the
 * reference, whose cylinder is solid (contact on its outside,
 * mphysics_cuda_impl.cuh:640-700); the contact law, constants and wall
 * conventions (ir = 0, wall point velocity in dv) are those of
 * cylinder_hertzian_contact_force_kernel (:720-774). */
typedef struct xdem_drum_info {
    xdem_double3 origin;
    xdem_double3 axis;
    double radius;
    double half_length;
    double omega;
    int32_t caps;        /* 1: flat end caps spinning with the shell; 0: open ends */
    int32_t pad_;
} xdem_drum_info;
#ifdef __cplusplus
static_assert(sizeof(xdem_drum_info) == 80, "xdem_drum_info layout");
#endif
/* Drum contact for every particle whose centre lies between the caps: one
 * shell contact (penetration r + rho - radius, rho = distance from the axis)
 * and, with caps, one cap contact (penetration r - (half_length - |a|)).
 * drum and cp are device pointers.  mf/mm (device, np entries, may be NULL)
 * receive each particle's reaction on the drum (normal and tangential force,
 * moment about `origin`); out_mf/out_mm (host) their sums. */
int32_t xdem_drum_contact_force_f64(int32_t tcm, const xdem_drum_info* drum, const double* pos,
                                    const double* vel, const double* omega, double* force,
                                    double* moment, const double* mass, uint32_t np,
                                    const xdem_device_contact_property* cp, xdem_double3* mf,
                                    xdem_double3* mm, xdem_double3* out_mf, xdem_double3* out_mm);
/* reductionD3 — mphysics_cuda_dec.cuh:157, def mphysics_cuda_def.cu:223-242.
 * Result written to host *out. */
int32_t xdem_reduction_d3(const xdem_double3* in, uint32_t n, xdem_double3* out);
/* cu_updatePolygonObjectData — mphysics_cuda_dec.cuh:296-298, def mphysics_cuda_def.cu:469-475 */
int32_t xdem_update_polygon_object_data_f64(const xdem_device_polygon_mass_info* dpmi,
                                            const double* vList, double* sphere,
                                            xdem_device_polygon_info* dpi, uint32_t ntriangle);

/* Contact pair list of the particle-particle predicate of cu_calculate_p2p
 * ((ri + rj) - |pj - pi| > 0, mphysics_cuda_impl.cuh:449-452): for each
 * particle i, counts[i] partners written to partners[i*max_per_particle ...]
 * in the kernel's traversal order (z, y, x neighbour cells, ascending sorted
 * position).  No counterpart in the reference: it exposes the pair set the
 * reference kernel iterates over, for bit-exact checks and tangential-history keys. */
int32_t xdem_contact_pairs_f64(const double* pos, const uint32_t* sorted_index,
                               const uint32_t* cstart, const uint32_t* cend, uint32_t np,
                               uint32_t max_per_particle, uint32_t* counts, uint32_t* partners);

/* ====================================================================
 * grid_base / neighborhood_cell — include/grid_base.h:7-64,
 * include/neighborhood_cell.h:6-24, source/neighborhood_cell.cpp:91-107.
 * An opaque grid object owning cell_id/body_id/sorted_id/cell_start/cell_end.
 * ==================================================================== */
typedef struct xdem_grid xdem_grid;
int32_t xdem_grid_create(xdem_grid** out);
int32_t xdem_grid_destroy(xdem_grid* g);
int32_t xdem_grid_set_world_origin(xdem_grid* g, double x, double y, double z);
int32_t xdem_grid_set_cell_size(xdem_grid* g, double cs);
int32_t xdem_grid_set_grid_size(xdem_grid* g, uint32_t gx, uint32_t gy, uint32_t gz);
int32_t xdem_grid_initialize(xdem_grid* g, uint32_t n);                 /* grid_base::initialize */
int32_t xdem_grid_detection(xdem_grid* g, const double* pos, const double* spos, uint32_t np,
                            uint32_t snp);                              /* neighborhood_cell::detection */
uint32_t xdem_grid_ncell(const xdem_grid* g);                           /* grid_base::nCell */
uint32_t* xdem_grid_sorted_id(xdem_grid* g);                            /* device pointers */
uint32_t* xdem_grid_cell_start(xdem_grid* g);
uint32_t* xdem_grid_cell_end(xdem_grid* g);
uint32_t* xdem_grid_cell_id(xdem_grid* g);
uint32_t* xdem_grid_body_id(xdem_grid* g);
/* grid_base::getCellNumber / getHash — source/grid_base.cpp:139-155 (host). */
int32_t xdem_grid_get_cell_number(const xdem_grid* g, double x, double y, double z, int32_t out[3]);
uint32_t xdem_grid_get_hash(const xdem_grid* g, const int32_t c3[3]);

/* ====================================================================
 * Single-scene simulation (dem_simulation::oneStepAnalysis,
 * source/simulation/dem_simulation.cpp:447-469) on the plain FP64 path.
 * ==================================================================== */
typedef struct xdem_sim xdem_sim;

typedef struct xdem_material {
    double density, youngs, poisson, shear;       /* shear modulus G; 0 => E / (2 (1 + nu)) */
} xdem_material;

typedef struct xdem_contact_pair_desc {
    double restitution, friction, rolling_friction, cohesion, stiffness_ratio;
    xdem_material wall;                           /* the j side (for particle-particle: same as particles) */
} xdem_contact_pair_desc;

typedef struct xdem_sim_desc {
    uint32_t np;                 /* particles */
    double dt;
    xdem_double3 gravity;
    xdem_double3 world_origin;
    double cell_size;            /* 0 => 2 * max radius (dem_simulation.cpp:261) */
    xdem_uint3 grid_size;        /* powers of two */
    bool rolling;                /* device_parameters.rollingCondition */
    int32_t tcm;                 /* 1 = DHS (live), 0 = HMCM */
    xdem_material particle;
    xdem_contact_pair_desc pp;   /* particle-particle */
} xdem_sim_desc;

int32_t xdem_sim_create(const xdem_sim_desc* d, const double* host_pos4, xdem_sim** out);
int32_t xdem_sim_destroy(xdem_sim* s);
/* Walls are applied in the order they are added (the reference applies
 * contacts in QMap name order, include/contactManager.h:88: add them in that order). */
int32_t xdem_sim_add_plane(xdem_sim* s, const xdem_device_plane_info* plane, const xdem_contact_pair_desc* c);
int32_t xdem_sim_add_cube(xdem_sim* s, const xdem_device_plane_info* planes6, const xdem_contact_pair_desc* c);
int32_t xdem_sim_add_cylinder(xdem_sim* s, const xdem_device_cylinder_info* cyl, const xdem_contact_pair_desc* c);
int32_t xdem_sim_set_polygons(xdem_sim* s, const xdem_device_polygon_info* tris, const double* sphere4,
                              uint32_t ntri, const xdem_device_polygon_mass_info* objs,
                              const xdem_contact_pair_desc* per_object, uint32_t nobj);
/* A rotating drum wall (xdem_drum_info), applied in order with the other walls. */
int32_t xdem_sim_add_drum(xdem_sim* s, const xdem_drum_info* drum, const xdem_contact_pair_desc* c);

/* Prescribed rigid motion of a plane or a polygon object, the reference's
 * geometry motion condition (source/xDynamicsSolver.cpp:176-184: before
 * every step whose time ct = k*dt satisfies ct >= start_time, the object is
 * moved by one step of its motion; contact_particles_plane.cpp:33-49 then
 * re-uploads the moved plane) with an angular velocity added for polygon
 * objects.  After n active steps the object has moved by velocity*n*dt and
 * turned by |omega|*n*dt about omega through its origin
 * (device_polygon_mass_info.origin): origin = origin0 + velocity*tau,
 * ep = q(omega, tau) * ep0 (Euler parameters, e0 first, Hamilton product),
 * triangle vertices = origin + A(ep) * local, where local = A(ep0)^T (P - origin0)
 * is taken when the motion is set.  end_time > start_time stops the motion
 * after end_time; otherwise it never stops.  While the motion is active the
 * object's vel/omega (device_polygon_mass_info) are velocity/omega, and zero
 * otherwise; a moving plane keeps the reference's stationary contact
 * velocity (mphysics_cuda_impl.cuh:600).  Planes take no rotation (omega must
 * be zero).  Setting a motion again replaces the previous one. */
typedef struct xdem_motion {
    xdem_double3 velocity;
    xdem_double3 omega;
    double start_time;
    double end_time;
} xdem_motion;
#ifdef __cplusplus
static_assert(sizeof(xdem_motion) == 64, "xdem_motion layout");
#endif
enum { XDEM_OBJECT_PLANE = 0, XDEM_OBJECT_POLYGON = 1 };
int32_t xdem_sim_set_motion(xdem_sim* s, int32_t kind, uint32_t index, const xdem_motion* m);
/* Host copies of the current triangles (ntri entries) and their circumspheres
 * (4*ntri doubles), either pointer may be NULL. */
int32_t xdem_sim_get_polygons(const xdem_sim* s, xdem_device_polygon_info* tris, double* sphere4);
/* Host copies of the polygon objects (nobj entries): pose, velocity and the
 * reaction force/moment of the last step's triangle contacts. */
int32_t xdem_sim_get_objects(const xdem_sim* s, xdem_device_polygon_mass_info* objs);
/* Hand-over from a multibody solver (contactManager::update ->
 * updatePolygonObjectData, contact_particles_polygonObjects.cpp:258-292):
 * the pose (origin, ep) and velocity (vel, omega) of object obj; its force and
 * moment fields are ignored.  The object's triangles are re-posed from their
 * local coordinates (taken when the triangles were set) before the next step.
 * The hand-over the other way is xdem_sim_get_objects (the reaction force and
 * moment of the last step, contact_particles_polygonObjects.cpp:469-479).
 * Refused for an object that has a prescribed motion. */
int32_t xdem_sim_set_object_state(xdem_sim* s, uint32_t obj,
                                  const xdem_device_polygon_mass_info* state);
/* Steps taken since creation (the current time is steps * dt). */
uint64_t xdem_sim_steps(const xdem_sim* s);

int32_t xdem_sim_step(xdem_sim* s, uint32_t nsteps);
/* Host copies of the state (any pointer may be NULL). */
int32_t xdem_sim_get_state(const xdem_sim* s, double* pos4, double* vel3, double* omega3,
                           double* acc3, double* alpha3, double* force3, double* moment3);
int32_t xdem_sim_set_state(xdem_sim* s, const double* pos4, const double* vel3, const double* omega3,
                           const double* acc3, const double* alpha3, const double* force3,
                           const double* moment3);
/* Device pointers of the state arrays (for callers that drive Set A themselves). */
int32_t xdem_sim_device_arrays(xdem_sim* s, double** pos, double** vel, double** omega, double** acc,
                               double** alpha, double** force, double** moment, double** mass,
                               double** iner);
uint32_t xdem_sim_np(const xdem_sim* s);
/* Real-time particle release (particleManager's real-time creation, run by
 * dem_simulation::initialize/oneStepAnalysis, dem_simulation.cpp:210-222,321-322,
 * 449-450): with every_steps > 0 the simulation starts with no particle active, and
 * at each step k (from 1) with (k - 1) % every_steps == 0 the next `group`
 * particles (in index order) join the simulated prefix; the others keep their
 * state untouched until then.  every_steps = 0 turns it off.  Before the first
 * step only.  xdem_sim_active gives the simulated count. */
int32_t xdem_sim_set_release(xdem_sim* s, uint32_t every_steps, uint32_t group);
/* The same release with a list of group sizes (particleManager::np_group, one
 * entry per generated layer or ring, particleManager.cpp:342-362,477; walked by
 * NextCreatingPerGroup, :132-139): the g-th release adds groups[g] particles,
 * and none once the list is exhausted.  every_steps > 0; the sizes sum to at
 * most xdem_sim_np.  Before the first step only. */
int32_t xdem_sim_set_release_groups(xdem_sim* s, uint32_t every_steps, const uint32_t* groups,
                                    uint32_t ngroups);
uint32_t xdem_sim_active(const xdem_sim* s);
/* Per-phase device time of xdem_sim_step, from HIP events recorded on the
 * library stream around each phase: 0 Verlet position, 1 hash, 2 sort,
 * 3 cell ranges, 4 particle-particle, 5 walls, 6 triangles, 7 Verlet velocity.
 * Enabling (or disabling) resets the totals. */
int32_t xdem_sim_enable_timing(xdem_sim* s, int32_t on);
int32_t xdem_sim_phase_times(xdem_sim* s, double* ms8, uint32_t* count8);
xdem_grid* xdem_sim_grid(xdem_sim* s);

#ifdef __cplusplus
}
#endif
#endif /* XDEM_H */
