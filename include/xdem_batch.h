/*
 * xdem_batch.h — the fused FP32c path: many independent DEM scenes advanced
 * together on one GPU (a segmented cell grid, scene id in the high part of
 * the cell key), with device-resident state and hipGraph-captured step blocks.
 *
 * Precision contract ("FP32c"):
 *   - positions are FP64 (accumulated in FP64 every step); the contact
 *     predicate (ri + rj) - |pj - pi| > 0 and the cell hash are evaluated in
 *     FP64 exactly as the reference (mphysics_cuda_impl.cuh:75-83, 449-452),
 *     so hashes, sorted indices, cell ranges and contact pair lists are
 *     bit-exact against the FP64 path on identical states;
 *   - radii, velocities, spins, accelerations and masses are FP32 (radii
 *     given in FP64 are rounded to FP32 on entry); the contact force law
 *     (Hertz normal, linear damping, Coulomb-capped tangential: the reference
 *     DHS model, mphysics_cuda_impl.cuh:379-401) runs in FP32.
 *
 * Semantics per scene are those of dem_simulation::oneStepAnalysis
 * (source/simulation/dem_simulation.cpp:447-469) with TCM = 1 (DHS): Verlet
 * position update, grid, particle-particle contacts in the reference's
 * traversal order (z, y, x neighbour cells; ascending sorted position), then
 * planes in the order given, then the Verlet velocity update with F reset to
 * m g.  No counterpart in the reference for the batching itself.
 */
#ifndef XDEM_BATCH_H
#define XDEM_BATCH_H

#include "xdem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xdem_batch xdem_batch;

typedef struct xdem_batch_scene_desc {
    uint32_t np;                   /* particles */
    double dt;
    xdem_double3 gravity;
    xdem_double3 world_origin;
    double cell_size;              /* > 0 */
    xdem_uint3 grid_size;          /* powers of two, product <= 2^24 */
    bool rolling;                  /* device_parameters.rollingCondition */
    xdem_material particle;
    xdem_contact_pair_desc pp;     /* particle-particle pair (pp.wall = particle material);
                                      pp.cohesion >= 0 adds cohesionForce
                                      (mphysics_cuda_impl.cuh:259-279) to the normal
                                      force (with tangential_history too) */
    xdem_contact_pair_desc wall;   /* particle-plane pair (wall.wall = wall material);
                                      wall contacts carry no cohesion term, as the
                                      FP64 entry points (xdem.h) */
    uint32_t nplane;
    const xdem_device_plane_info* planes;  /* host array of nplane planes, applied in order */
    int32_t tangential_history;    /* 0: the reference DHS law (memoryless tangential force);
                                      1: DHS with a tangential spring per contact (below) */
} xdem_batch_scene_desc;

/* Tangential-spring history (tangential_history = 1; no counterpart in the
 * reference).  Per particle up to XDEM_BATCH_HISTORY_SLOTS contacts, in
 * evaluation order (particle contacts in the reference traversal order, then
 * planes in order, then drums in order: shell, then the cap on the particle's
 * side), each a key (partner particle index within the scene,
 * XDEM_BATCH_WALL_KEY | plane index, or XDEM_BATCH_DRUM_KEY | drum index << 2
 * | side with side 0 the shell, 1 the cap at +half_length, 2 the cap at
 * -half_length) and a spring xi (FP32).  Per step: xi is
 * projected onto the current tangent plane and advanced by dt e (e the
 * tangential relative velocity), Ft = ks xi + vs e; beyond the Coulomb cap
 * mu |Fn| the force is capped and xi reset to (Ft - vs e) / ks.  A contact
 * that ends drops its spring.  Contacts beyond the slots are evaluated with a
 * zero spring and counted (xdem_batch_history_overflow). */
#define XDEM_BATCH_HISTORY_SLOTS 16
#define XDEM_BATCH_WALL_KEY 0x80000000u
#define XDEM_BATCH_DRUM_KEY 0xC0000000u

/* Walls beyond the planes, and triangle meshes, for a scene that has not
 * stepped yet (no counterpart in the reference's batched form; the contact
 * rules are those of the FP64 entry points in xdem.h).  Polygon objects are
 * not available with tangential_history; drums are (their contacts keep
 * springs as above).
 *
 * xdem_batch_add_drum: a rotating drum (xdem_drum_info, as
 * xdem_drum_contact_force_f64), applied after the scene's planes, in the order
 * added; wall = the particle-drum pair.
 *
 * xdem_batch_add_polygon_object: a triangle mesh (ntri triangles in world
 * coordinates at the scene's step 0; .id ignored) attached to a frame at
 * origin/ep (Euler parameters, e0 first), moving with motion->velocity and
 * motion->omega about origin at the steps whose time k dt (the scene's step k,
 * from 1) is >= start_time and, when end_time > start_time, <= end_time, as
 * the simulation object's motion (motion NULL: static).  Contacts follow
 * particle_polygonObject_collision_kernel (mphysics_cuda_impl.cuh:923-1021):
 * every triangle within reach counts, normal = -(Q-P)x(R-P)/|.|.  The
 * triangles are binned once in the object's frame; the scene's cell size must
 * be at least its largest radius plus the largest circumradius (refused
 * otherwise), which makes the search complete.  Reactions:
 * xdem_batch_object_reactions. */
int32_t xdem_batch_add_drum(xdem_batch* b, uint32_t scene_id, const xdem_drum_info* drum,
                            const xdem_contact_pair_desc* wall);
int32_t xdem_batch_add_polygon_object(xdem_batch* b, uint32_t scene_id,
                                      const xdem_device_polygon_info* tris, uint32_t ntri,
                                      const xdem_double3* origin, const xdem_double4* ep,
                                      const xdem_motion* motion,
                                      const xdem_contact_pair_desc* wall);

/* Capacities are fixed at creation (HBM is reserved up front): at most 2^27
 * particles and 2^29 cells per batch (the contact kernel addresses its arrays
 * with 32-bit byte offsets); larger runs use several batches. */
int32_t xdem_batch_create(uint32_t max_particles, uint32_t max_scenes, uint64_t max_cells,
                          xdem_batch** out);
/* Device bytes a batch of these capacities holds: what xdem_batch_create
 * allocates, its scene and launch tables at their largest, with
 * XDEM_BATCH_FOOTPRINT_HISTORY the history buffers the first such scene adds,
 * with XDEM_BATCH_FOOTPRINT_LISTS the candidate lists of xdem_batch_set_lists
 * (the parity probe's buffers are not counted).  Host arithmetic only (no GPU
 * needed); admission control budgets HBM with it. */
#define XDEM_BATCH_FOOTPRINT_HISTORY 1
#define XDEM_BATCH_FOOTPRINT_LISTS 2
uint64_t xdem_batch_device_bytes(uint32_t max_particles, uint32_t max_scenes, uint64_t max_cells,
                                 int32_t options);
/* Candidate lists (no counterpart in the reference): the contact kernel keeps,
 * per particle, the neighbours within ri + rj + skin (skin = skin_fraction x
 * the scene's cell size; 0 keeps the current fraction, 0.02 at creation) in
 * the reference traversal order and evaluates the contact predicate on them
 * instead of walking the 27 cells.  A particle walks its cells again (and
 * rebuilds its list) at the first step of every xdem_batch_step call, when
 * its scene's accumulated displacement bound reaches half the skin, when more
 * than 1024 particles of the batch changed cell in the last step, when it moved
 * in the sorted order, or when a particle changed cell within one cell of its
 * own; contacts and their order, hence the results, are bit-identical to the
 * walk.  Plane-only batches (no history, drums or meshes); the lists take
 * (32 + 1) x 4 bytes per particle of capacity and 2 bytes per cell. */
int32_t xdem_batch_set_lists(xdem_batch* b, int32_t on, float skin_fraction);
/* Scenes whose lists were all rebuilt for the last step (the rest reused
 * theirs, particle by particle as above). */
int32_t xdem_batch_list_rebuilds(xdem_batch* b, uint32_t* scenes_rebuilding);
/* The re-sort between steps: with at most 1024 particles changed cell (on, the
 * default) only the windows of the sorted order those particles shift are
 * rewritten, in one kernel; off: the full merge every step.  The results are
 * identical. */
int32_t xdem_batch_set_sort_patch(xdem_batch* b, int32_t on);
/* After xdem_batch_prepare: capture and instantiate the graph that replays an
 * nsteps block (both buffer parities when nsteps is odd) without running it,
 * so the first xdem_batch_step(b, nsteps) of a timed run only launches.  A
 * change of launch shape, walk, lists or scenes drops the graphs.  No-op with
 * graphs off. */
int32_t xdem_batch_capture(xdem_batch* b, uint32_t nsteps);
int32_t xdem_batch_destroy(xdem_batch* b);
/* Add a scene (host arrays, reference layout: pos4 = x,y,z,r; vel3/omega3 may be
 * NULL for zero).  Takes effect at the next xdem_batch_step.  Fails with
 * XDEM_ERR_CAPACITY when particles, scenes or cells would exceed the capacity. */
int32_t xdem_batch_add_scene(xdem_batch* b, const xdem_batch_scene_desc* d, const double* pos4,
                             const double* vel3, const double* omega3, uint32_t* scene_id);
/* As xdem_batch_add_scene, also restoring the accelerations (acc3, alpha3 may
 * be NULL).  With the arrays read back by xdem_batch_get_state this resumes a
 * scene bit-exactly (checkpoint/resume; the reference's resume drops the
 * accelerations, source/simulation/dem_simulation.cpp:585-593). */
int32_t xdem_batch_add_scene_state(xdem_batch* b, const xdem_batch_scene_desc* d,
                                   const double* pos4, const double* vel3, const double* omega3,
                                   const double* acc3, const double* alpha3, uint32_t* scene_id);
/* As xdem_batch_add_scene_state, also restoring the tangential-spring history
 * (hist_count[np], hist_key[np * SLOTS], hist_xi[np * SLOTS * 3], original
 * particle order; all NULL for none) as read back by xdem_batch_get_history. */
int32_t xdem_batch_add_scene_history(xdem_batch* b, const xdem_batch_scene_desc* d,
                                     const double* pos4, const double* vel3, const double* omega3,
                                     const double* acc3, const double* alpha3,
                                     const uint32_t* hist_count, const uint32_t* hist_key,
                                     const double* hist_xi, uint32_t* scene_id);
/* Remove a scene (its slot and memory are reused).  Takes effect at the next step. */
int32_t xdem_batch_remove_scene(xdem_batch* b, uint32_t scene_id);
/* Advance every active scene by nsteps (one step block).  When graphs are on
 * (default), the block is captured once per (layout, nsteps) and replayed. */
int32_t xdem_batch_step(xdem_batch* b, uint32_t nsteps);
int32_t xdem_batch_set_graphs(xdem_batch* b, int32_t on);
/* Block size (64, 128, 192 or 256) of the contact kernel; 0 = default (256). */
int32_t xdem_batch_set_block(xdem_batch* b, uint32_t threads);
/* The contact kernel's launch shape without autotuning: block size (as
 * xdem_batch_set_block), neighbour candidates loaded together (1, 2 or 4) and
 * contact records loaded one contact ahead (0/1; ignored while a scene has
 * tangential history).  Results do not depend on it. */
int32_t xdem_batch_set_launch(xdem_batch* b, uint32_t threads, uint32_t unroll, int32_t load_ahead);
/* How a thread walks its neighbour rows (the 27-cell loop of calculate_p2p_kernel,
 * mphysics_cuda_impl.cuh:435-482, in the same order either way): per_lane = 1
 * (default) walks its own non-empty rows as one stream, 0 walks row by row in
 * step with the other lanes of its wave.  Results do not depend on it. */
int32_t xdem_batch_set_walk(xdem_batch* b, int32_t per_lane);
/* Autotuner: for each launch shape — neighbour search (the walk row by row, the
 * per-lane walk, or the per-lane walk with candidate lists, xdem_batch_set_walk /
 * xdem_batch_set_lists), block sizes 64, 128 and 256, 1, 2 or 4 neighbour
 * candidates loaded together, contact records loaded one contact ahead or not —
 * times a whole graph-replayed block of steps_per_trial steps (rounded up to
 * even; 0 means 20) after a warm block of the same length (HIP events on the
 * batch stream: contacts, Verlet and re-sort), and keeps the fastest.  The
 * scenes advance 54 * (2 * steps_per_trial + 1) steps at most.  *chosen = block
 * size + 1000 * candidates-per-load-batch + 10000 * (1 if contact records are
 * loaded ahead) + 100000 * (1 for the per-lane walk) + 1000000 * (1 with
 * candidate lists); ms_per_candidate (54 entries: search-major, then block size,
 * then candidates-per-batch, then load-ahead off/on; may be NULL) receives the
 * measured time per step (-1 for shapes not used here: load-ahead while a scene
 * has history on or drums/meshes, and candidate lists then too).  Results do
 * not depend on the choice (bit-identical). */
int32_t xdem_batch_autotune(xdem_batch* b, uint32_t steps_per_trial, uint32_t* chosen,
                            double* ms_per_candidate);
int32_t xdem_batch_synchronize(xdem_batch* b);
/* State of one scene in its original particle order (host arrays, reference
 * layout; any pointer may be NULL).  Positions are FP64; the others are the
 * FP32 values widened. */
int32_t xdem_batch_get_state(xdem_batch* b, uint32_t scene_id, double* pos4, double* vel3,
                             double* omega3, double* acc3, double* alpha3);
/* Tangential-spring history of one scene in original particle order (key and
 * xi may be NULL; unused slots read 0).  Scenes without history read count 0. */
int32_t xdem_batch_get_history(xdem_batch* b, uint32_t scene_id, uint32_t* count, uint32_t* key,
                               double* xi);
/* Contacts found with every history slot of their particle taken, summed over
 * all steps since creation. */
int32_t xdem_batch_history_overflow(xdem_batch* b, uint32_t* overflow);
uint32_t xdem_batch_num_particles(const xdem_batch* b);
/* Particles of one scene (0 when there is no such scene). */
uint32_t xdem_batch_scene_particles(const xdem_batch* b, uint32_t scene_id);
uint32_t xdem_batch_num_scenes(const xdem_batch* b);
/* Kinetic energy (translational + rotational) and linear momentum of one
 * scene, reduced on the device in FP64 (deterministic order). */
int32_t xdem_batch_energy(xdem_batch* b, uint32_t scene_id, double* kinetic, double momentum[3]);
/* Contact metrics of a scene at its current positions (after
 * xdem_batch_prepare): particle-particle contacts, each pair once (the
 * reference predicate (ri + rj) - |pj - pi| > 0 in FP64 over the 27 cells of
 * the walk), and the largest overlap (0 without contacts).  No counterpart in
 * the reference, which has no contact counters; for per-block logs. */
int32_t xdem_batch_contact_metrics(xdem_batch* b, uint32_t scene_id, uint64_t* contact_pairs,
                                   double* max_overlap);
/* Reactions of the triangle-mesh objects, the batched counterpart of the
 * per-object force/moment the reference accumulates in dpmi[] during the
 * triangle contact kernel (particle_polygonObject_collision_kernel,
 * include/mphysics_cuda_impl.cuh:1008-1009; read back by
 * contact_particles_polygonObjects.cpp:470-479): the reaction force
 * -(sum of Fn + Ft over the object's triangle contacts) and its moment about the
 * object origin.  When on, the last step of every xdem_batch_step evaluates
 * them, reduced in a fixed order (identical on every run).  Toggling rebuilds
 * the captured step graphs. */
int32_t xdem_batch_enable_reactions(xdem_batch* b, int32_t on);
/* The reactions of the last step of the scene's objects, in the order they were
 * added; *nobj = the scene's object count (force/moment may be NULL to query it). */
int32_t xdem_batch_object_reactions(xdem_batch* b, uint32_t scene_id, xdem_double3* force,
                                    xdem_double3* moment, uint32_t cap, uint32_t* nobj);
/* The same for the scene's drum walls (order of xdem_batch_add_drum): minus the
 * summed drum contact forces and their moment about the drum origin (the drive
 * torque is its component along the axis), as xdem_drum_contact_force_f64's
 * out_mf/out_mm; evaluated with the object reactions when those are on.
 * Refused for a scene with tangential_history (that evaluation has no springs). */
int32_t xdem_batch_drum_reactions(xdem_batch* b, uint32_t scene_id, xdem_double3* force,
                                  xdem_double3* moment, uint32_t cap, uint32_t* ndrum);
/* Per-phase device times (HIP events), phases: 0 position update + hash,
 * 1 contact + Verlet, 2 sort (mover detection), 3 sort (mover ordering),
 * 4 sort (merge), 5 cell ranges.  Enabling resets the totals. */
int32_t xdem_batch_enable_timing(xdem_batch* b, int32_t on);
int32_t xdem_batch_phase_times(xdem_batch* b, double* ms6, uint32_t* count6);
/* Number of particles that changed cell at the last sort (all scenes). */
int32_t xdem_batch_last_movers(xdem_batch* b, uint32_t* movers);

/* Parity hooks (no counterpart in the reference).  After xdem_batch_prepare,
 * the batch holds the sorted grid of the current positions (exactly as the
 * reference computes it from the positions given to add_scene, i.e. before
 * any Verlet update) and xdem_batch_probe evaluates the contact step on that
 * grid without advancing: per particle (original order of `scene_id`) it
 * returns the particle-particle plus wall force and torque (FP32 widened),
 * the contact partner ids in traversal order (up to max_per), and the sorted
 * cell keys / sorted ids / cell ranges of that scene. */
int32_t xdem_batch_prepare(xdem_batch* b);
int32_t xdem_batch_probe(xdem_batch* b, uint32_t scene_id, double* force3, double* moment3,
                         uint32_t max_per, uint32_t* counts, uint32_t* partners,
                         uint32_t* sorted_keys, uint32_t* sorted_ids, uint32_t* cstart,
                         uint32_t* cend);

#ifdef __cplusplus
}
#endif
#endif
