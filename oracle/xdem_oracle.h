/*
 * xdem_oracle.h — CPU restatement (FP64) of the reference DEM contact step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (libxdem.so) never calls it.
 *
 * Two semantics (the reference's CPU path and its GPU kernels differ):
 *   ORC_GPU  (default) follows include/mphysics_cuda_impl.cuh:65-1021,1898-1956
 *   ORC_HOST follows the serial CPU path: source/neighborhood_cell.cpp:31-84,
 *            source/contactManager.cpp:208-277, source/collision/contact.cpp:20-46,
 *            source/collision/contact_particles_{particles,plane,cube}.cpp,
 *            source/velocity_verlet.cpp:18-67.
 * Differences between the two are listed per function in xdem_oracle.cpp.
 *
 * Parity status per function (what pins it) is recorded in the header comment
 * of xdem_oracle.cpp and checked by tests/test_oracle_*.py.
 */
#ifndef XDEM_ORACLE_H
#define XDEM_ORACLE_H
#include "../include/xdem.h"
#include "../include/xdem_setb.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_GPU = 0, ORC_HOST = 1 };
enum { ORC_WALL_PLANE = 0, ORC_WALL_CYLINDER = 1, ORC_WALL_DRUM = 2 };

/* One wall contact in application order. */
typedef struct orc_wall {
    int32_t kind;          /* ORC_WALL_PLANE / ORC_WALL_CYLINDER / ORC_WALL_DRUM */
    int32_t index;         /* into planes[], cylinders[] or drums[] */
    int32_t cube_member;   /* 1: plane of a cube (CPU path accumulates; a lone plane assigns) */
    int32_t pad_;
    xdem_device_contact_property cp;
} orc_wall;

/* A prescribed motion (xdem_motion) of plane or polygon object `index`
 * (kind XDEM_OBJECT_PLANE / XDEM_OBJECT_POLYGON); n_act counts active steps. */
typedef struct orc_motion {
    int32_t kind;
    uint32_t index;
    xdem_motion m;
    uint64_t n_act;
    int32_t active;
    int32_t pad_;
} orc_motion;

/* A whole single-scene state, host memory, reference layout. */
typedef struct orc_world {
    xdem_device_parameters prm;  /* np, grid, dt, gravity, rolling ... */
    int32_t tcm;
    int32_t semantics;           /* ORC_GPU / ORC_HOST */
    uint32_t np;
    double* pos;    /* 4*np */
    double* vel;    /* 3*np */
    double* omega;
    double* acc;
    double* alpha;
    double* force;
    double* moment;
    double* mass;   /* np */
    double* iner;
    xdem_device_contact_property cp_pp;
    uint32_t nwall;
    const orc_wall* walls;
    xdem_device_plane_info* planes;   /* moving planes are rewritten every step */
    const xdem_device_cylinder_info* cylinders;
    /* triangles (polygon objects); rewritten every step when an object moves */
    uint32_t ntri;
    xdem_device_polygon_info* tri;
    double* tri_sphere;          /* 4*ntri circumspheres */
    uint32_t nobj;
    xdem_device_polygon_mass_info* obj;
    const xdem_device_contact_property* cp_obj;
    /* scratch owned by the caller: n = np + ntri entries / ncell entries */
    uint32_t* hash;
    uint32_t* index;
    uint32_t* sorted;
    uint32_t* cstart;
    uint32_t* cend;
    /* drums and prescribed motion */
    const xdem_drum_info* drums;
    uint32_t nmotion;
    orc_motion* motions;
    const xdem_device_plane_info* planes0;       /* planes as added */
    const xdem_device_polygon_mass_info* obj0;   /* object poses as set */
    const double* vlist;                         /* 9*ntri local vertex coordinates */
    uint64_t step;                               /* steps taken */
    /* real-time release (dem_simulation::oneStepAnalysis, dem_simulation.cpp:449-450):
     * with release_every != 0 only the first np of np_total particles are simulated, and
     * at step k (from 1) with (k - 1) % release_every == 0 the next group joins:
     * release_group particles, or with nrelease_groups != 0 the entry release_next
     * of release_groups (none once the list is exhausted,
     * particleManager::NextCreatingPerGroup, particleManager.cpp:132-139) */
    uint32_t np_total;
    uint32_t release_every;
    uint32_t release_group;
    uint32_t nrelease_groups;
    uint32_t release_next;
    const uint32_t* release_groups;
} orc_world;

/* grid_base::getCellNumber / calcGridPos + calcGridHash */
void orc_cell_of(const xdem_device_parameters* p, int semantics, double x, double y, double z,
                 int32_t out[3]);
uint32_t orc_hash_of(const xdem_device_parameters* p, const int32_t c[3]);
/* K3 / K4: hash[off+i], index[off+i] = off+i for i < n */
void orc_hash_and_index(const xdem_device_parameters* p, int semantics, const double* pos4,
                        uint32_t n, uint32_t off, uint32_t* hash, uint32_t* index);
/* K5-K7: stable sort by hash, cell ranges over ncell, sorted_index */
void orc_reorder_and_find_cell_start(uint32_t* hash, uint32_t* index, uint32_t n, uint32_t ncell,
                                     uint32_t* cstart, uint32_t* cend, uint32_t* sorted);
/* cohesionForce (mphysics_cuda_impl.cuh:259-279) and contact::cohesionForce
 * (source/collision/contact.cpp:6-18) */
double orc_cohesion_force(double ri, double rj, double Ei, double Ej, double pri, double prj,
                          double coh, double Fn);
double orc_host_cohesion(double coh, double coh_r, double coh_e, double Fn);
/* From now on every contact force Fn + Ft the oracle evaluates (particle,
 * plane, cylinder, drum, triangle, history twin) adds |Fn + Ft| to fmag[id]
 * of its particle and the magnitudes of the terms of the contact law,
 * |fsn| + |fca| + (|vn| + |vs|) V + |Ft| (V: summed magnitudes of the
 * velocities entering the relative velocity), to tmag[id]; NULL stops. */
void orc_set_force_magnitudes(double* fmag, double* tmag);
/* getConstant (mphysics_cuda_impl.cuh:214-256) */
void orc_force_constant(int32_t tcm, double ir, double jr, double im, double jm,
                        const xdem_device_contact_property* cp, xdem_device_force_constant* out);
/* K8 (GPU) / hostCollision p2p part (HOST, adds into force/moment).  If pairs
 * is non-null, the (i,k) contact pairs are appended in traversal order
 * (at most max_pairs); the total count is returned. */
uint64_t orc_p2p(const xdem_device_parameters* p, int32_t tcm, int semantics, const double* pos4,
                 const double* vel, const double* omega, double* force, double* moment,
                 const double* mass, const uint32_t* sorted, const uint32_t* cstart,
                 const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np,
                 uint32_t* pairs, uint64_t max_pairs);
/* Brute-force O(N^2) pair predicate, pairs (i,k) i!=k, ordered by i then k. */
uint64_t orc_pairs_bruteforce(const double* pos4, uint32_t np, uint32_t* pairs, uint64_t max_pairs);
/* particle_plane_contact_detection (returns cdist, writes unit) */
double orc_plane_detect(const xdem_device_plane_info* pe, int semantics, const double xp[3],
                        const double wp[3], double u[3], double r);
/* K9 on one plane (GPU: force += ...) */
void orc_plane(const xdem_device_parameters* p, int32_t tcm, const xdem_device_plane_info* pl,
               const double* pos4, const double* vel, const double* omega, double* force,
               double* moment, const double* mass, uint32_t np,
               const xdem_device_contact_property* cp);
/* cylinder detection (returns float-truncated cdist as the reference) */
double orc_cylinder_detect(const xdem_device_cylinder_info* cy, const double pt[4], double u[3],
                           double cpt[3]);
void orc_cylinder(const xdem_device_parameters* p, int32_t tcm, const xdem_device_cylinder_info* cy,
                  const double* pos4, const double* vel, const double* omega, double* force,
                  double* moment, const double* mass, uint32_t np,
                  const xdem_device_contact_property* cp, const double* mpos3, double* mf3,
                  double* mm3);
/* closestPtPointTriangle (mphysics_cuda_impl.cuh:837-899) */
void orc_closest_pt_triangle(const xdem_device_polygon_info* t, const double p[3], double out[3]);
/* K11; reaction sums (per object, zeroed first) go to obj[].force/moment when obj_reaction != 0 */
void orc_polygon(const xdem_device_parameters* p, int32_t tcm, const xdem_device_polygon_info* dpi,
                 xdem_device_polygon_mass_info* dpmi, uint32_t nobj, int obj_reaction,
                 const double* pos4, const double* vel, const double* omega, double* force,
                 double* moment, const double* mass, const uint32_t* sorted,
                 const uint32_t* cstart, const uint32_t* cend,
                 const xdem_device_contact_property* cp, uint32_t np);
/* K1 / K2 (GPU) or velocity_verlet CPU branch (HOST) */
void orc_vv_position(const xdem_device_parameters* p, int semantics, double* pos4, const double* vel,
                     const double* acc, uint32_t np);
void orc_vv_velocity(const xdem_device_parameters* p, int semantics, double* vel, double* acc,
                     double* omega, double* alpha, double* force, double* moment,
                     const double* mass, const double* iner, uint32_t np);
/* reductionD3: same block-tree order as reduce6<double3,512> + host block sum */
void orc_reduction_d3(const double* in3, uint32_t n, double out[3]);
/* updatePolygonObjectData_kernel (mphysics_cuda_impl.cuh:1911-1956) */
/* Rotating drum contact (xdem_drum_contact_force_f64); mf3/mm3 may be NULL. */
void orc_drum(const xdem_device_parameters* p, int32_t tcm, const xdem_drum_info* dr,
              const double* pos4, const double* vel, const double* omega, double* force,
              double* moment, const double* mass, uint32_t np,
              const xdem_device_contact_property* cp, double* mf3, double* mm3);
/* Pose after tau seconds of motion m from (o0, ep0). */
void orc_motion_pose(const xdem_double3* o0, const xdem_double4* ep0, const xdem_motion* m,
                     double tau, xdem_double3* o, xdem_double4* ep);
/* Local vertex coordinates A(ep0)^T (P - origin0), 9 per triangle. */
void orc_local_vertices(const xdem_device_polygon_info* tris, uint32_t ntri,
                        const xdem_device_polygon_mass_info* objs, double* vlist);
void orc_update_polygon(const xdem_device_polygon_mass_info* dpmi, const double* vList,
                        double* sphere4, xdem_device_polygon_info* dpi, uint32_t ntri);
/* numeric::utility::calculate_center_of_triangle (source/numeric_utility.cpp:20-45) */
void orc_circumcenter(const double P[3], const double Q[3], const double R[3], double out[3]);
/* dem_simulation::oneStepAnalysis (dem_simulation.cpp:447-469), nsteps times */
void orc_step(orc_world* w, uint32_t nsteps);
/* Parallel (OpenMP over particles) variant of orc_step for the multi-core
 * CPU baseline; identical results (each particle is independent). */
void orc_step_omp(orc_world* w, uint32_t nsteps, int32_t nthreads);

/* Tangential-spring history (DHS with a per-contact tangential spring).  No
 * counterpart in the reference, whose DHS tangential force is memoryless
 * (mphysics_cuda_impl.cuh:379-401, contact.cpp:20-46).  Per particle, up to
 * `slots` contacts in evaluation order (particle contacts in the traversal
 * order of calculate_p2p_kernel, then planes in application order, then
 * drums in order: shell, then the cap on the particle's side); key = partner
 * particle id, ORC_WALL_KEY | wall index for a plane, or ORC_DRUM_KEY |
 * (drum index << 2 | side) for a drum (side 0 shell, 1 cap at +half_length,
 * 2 cap at -half_length); xi = spring (3 doubles).
 * Law (own definition): Fn as DHS (particle pairs with the cohesionForce term);
 * e = dv - (dv.n)n; xi <- xi - (xi.n)n + dt e;
 * Ft = ks xi + vs e; if |Ft| > mu |Fn| then Ft <- mu |Fn| Ft / |Ft| and
 * xi <- (Ft - vs e) / ks (the spring that reproduces the capped force);
 * M = (ri n) x Ft (walls: ri = 0).  With xi = 0 on entry this
 * is the reference DHS tangential force whenever |e| != 0. */
#define ORC_WALL_KEY 0x80000000u
#define ORC_DRUM_KEY 0xC0000000u
typedef struct orc_hist {
    uint32_t slots;
    uint32_t pad_;
    uint32_t* count;   /* np */
    uint32_t* key;     /* np * slots */
    double* xi;        /* np * slots * 3 */
    uint64_t overflow; /* contacts found with all slots taken (not kept) */
} orc_hist;

/* Contacts of the current state with history: force = gravity-free sum of
 * particle and plane contacts, moment likewise (both overwritten), `out`
 * receives the new history.  Planes only among the walls (TCM 1, GPU
 * semantics).  The grid (w->sorted/cstart/cend) must be current. */
void orc_hist_contacts(const orc_world* w, const orc_hist* in, orc_hist* out, double* force,
                       double* moment);
/* Rebuild w's grid from its positions (hash, sort, cell ranges). */
void orc_detect(orc_world* w);
/* One step with history (GPU semantics, TCM 1): Verlet position, grid,
 * contacts with history (in -> out), Verlet velocity with F = m g + contacts. */
void orc_hist_step(orc_world* w, const orc_hist* in, orc_hist* out);

/* The older entry points (include/cu_dem_dec.cuh), restated in
 * xdem_oracle_b.cpp; arguments as the xdem_b_* functions of xdem_setb.h,
 * host arrays. */
void orcb_vv_position(const xdem_b_device_parameters* p, double* pos4, const double* vel4,
                      const double* acc4, uint32_t np);
void orcb_vv_velocity(const xdem_b_device_parameters* p, double* vel4, double* acc4,
                      double* omega4, double* alpha4, const double* force3,
                      const double* moment3, uint32_t np);
void orcb_mergedata(double* data4, const double* pos4, uint32_t np, const double* vertice3,
                    uint32_t snp);
void orcb_hash_and_index(const xdem_b_device_parameters* p, uint32_t* hash, uint32_t* index,
                         const double* pos4, uint32_t np);
void orcb_reorder(const xdem_b_device_parameters* p, const uint32_t* hash, const uint32_t* index,
                  uint32_t* cstart, uint32_t* cend, uint32_t* sorted, uint32_t np, uint32_t nc);
void orcb_p2p(const xdem_b_device_parameters* p, const double* pos4, const double* vel4,
              const double* acc4, const double* omega4, double* force3, double* moment3,
              const uint32_t* sorted, const uint32_t* cstart, const uint32_t* cend, uint32_t np);
void orcb_cube(const xdem_b_device_parameters* p, const xdem_b_plane_info* planes, double kn,
               double vn, double ks, double vs, double mu, const double* pos4, const double* vel4,
               const double* omega4, double* force3, double* moment3, uint32_t np);
void orcb_shape(const xdem_b_device_parameters* p, const xdem_b_mass_info* mi,
                const xdem_b_polygon* polys, const uint32_t* id_set, const uint32_t* poly_start,
                const uint32_t* poly_end, double kn, double vn, double ks, double vs, double mu,
                bool* is_line, const double* pos4, const double* vel4, const double* omega4,
                double* force3, double* moment3, uint32_t np, const uint32_t* sorted,
                const uint32_t* cstart, const uint32_t* cend, double out_bf[3]);
void orcb_update_polygons_position(const double* mpos3, const double* A, const double* local3,
                                   uint32_t nv, double* pos3);
void orcb_update_polygons_information(xdem_b_polygon* polys, const uint32_t* indice3,
                                      const double* pos3, uint32_t ni);

#ifdef __cplusplus
}
#endif
#endif
