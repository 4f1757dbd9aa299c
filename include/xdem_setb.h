/*
 * xdem_setb.h — the older DEM entry points of the reference
 * (include/cu_dem_dec.cuh:95-113, source/cu_dem_def.cu:25-244, kernels in
 * include/cu_dem_impl.cuh), as used by its parSIM generation
 * (source/cell_grid.cpp:208-227, source/force.cpp:144-156,
 * source/geometry.cpp:262,786,816-817, source/Verlet_integrator.cpp:56,60).
 *
 * Same conventions as xdem.h: device pointers in, int32_t status out
 * (0 = ok, message from xdem_last_error()), the parameter block held by the
 * library (xdem_b_set_symbolic_parameter), launches on the xdem_set_stream
 * stream.  Layout of this set: pos, vel, acc, omega, alpha are double4 per
 * particle with acc.w = mass and alpha.w = inertia; force and moment are
 * packed double3.  FP64 throughout.
 *
 * Semantics follow the reference kernels, including their launch bounds:
 * a kernel launched for `np` items processes the ids below
 * min(T(np), bound), with T(n) = ceil(n / min(256, n)) * min(256, n) the
 * reference's thread count and `bound` the parameter block's np (np + nsp
 * for the hash and the cell ranges).  Deliberate departures (defects not
 * reproduced): no per-call device allocation in
 * xdem_b_shape_hertzian_contact_force; reduction lanes beyond the particle
 * count read zero; np == 0 is a no-op (the reference divides by zero).
 */
#ifndef XDEM_SETB_H
#define XDEM_SETB_H

#include "xdem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cu_dem_dec.cuh:41-60 (144 B) */
typedef struct xdem_b_device_parameters {
    uint32_t np;
    uint32_t nsp;                  /* shape vertices appended after the particles */
    uint32_t ncell;
    xdem_uint3 grid_size;
    double dt;
    double half2dt;
    double cell_size;
    xdem_double3 gravity;
    xdem_double3 world_origin;
    double kn, vn, ks, vs, mu;     /* particle-particle constants */
    double cohesive;
} xdem_b_device_parameters;

/* cu_dem_dec.cuh:62-71 (160 B) */
typedef struct xdem_b_plane_info {
    double l1, l2;
    xdem_double3 u1, u2, uw, xw, pa, pb;
} xdem_b_plane_info;

/* cu_dem_dec.cuh:73-81 (144 B) */
typedef struct xdem_b_polygon {
    xdem_double3 P, Q, R, V, W, N;
} xdem_b_polygon;

/* cu_dem_dec.cuh:83-87 (48 B) */
typedef struct xdem_b_mass_info {
    xdem_double3 pos;
    xdem_double3 vel;
} xdem_b_mass_info;

/* setSymbolicParameter — cu_dem_def.cu:25-28 */
int32_t xdem_b_set_symbolic_parameter(const xdem_b_device_parameters* p);
int32_t xdem_b_get_symbolic_parameter(xdem_b_device_parameters* p);

/* vv_update_position — cu_dem_def.cu:30-37, kernel cu_dem_impl.cuh:93-107 */
int32_t xdem_b_vv_update_position(double* pos4, const double* vel4, const double* acc4,
                                  uint32_t np);
/* vv_update_velocity — cu_dem_def.cu:39-50, kernel cu_dem_impl.cuh:109-140 */
int32_t xdem_b_vv_update_velocity(double* vel4, double* acc4, double* omega4, double* alpha4,
                                  const double* force3, const double* moment3, uint32_t np);
/* cu_mergedata — cu_dem_def.cu:51-55, kernel cu_dem_impl.cuh:782-796:
 * data[i] = pos[i] (i < np), data[np + k] = (vertice[k], -1). */
int32_t xdem_b_mergedata(double* data4, const double* pos4, uint32_t np, const double* vertice3,
                         uint32_t snp);
/* cu_calculateHashAndIndex — cu_dem_def.cu:57-69 (kernel cu_dem_impl.cuh:142-154): hash
 * and index, then the
 * stable sort of (hash[0..np), index) by hash. */
int32_t xdem_b_calculate_hash_and_index(uint32_t* hash, uint32_t* index, const double* pos4,
                                        uint32_t np);
/* cu_reorderDataAndFindCellStart — cu_dem_def.cu:71-90, kernel
 * cu_dem_impl.cuh:156-196: cstart cleared to 0xFFFFFFFF (cend is not). */
int32_t xdem_b_reorder_data_and_find_cell_start(const uint32_t* hash, const uint32_t* index,
                                                uint32_t* cstart, uint32_t* cend,
                                                uint32_t* sorted_index, uint32_t np, uint32_t nc);
/* cu_calculate_p2p — cu_dem_def.cu:92-107, kernels cu_dem_impl.cuh:198-327 (calForce,
 * calculate_p2p_kernel):
 * force[i] += acc[i].w g + sum of normal forces (+ cohesion), moment[i] +=
 * sum of (ri n) x Ft; constants from the parameter block.  cRun is unused
 * by the reference kernel. */
int32_t xdem_b_calculate_p2p(const double* pos4, const double* vel4, const double* acc4,
                             const double* omega4, const double* alpha4, double* force3,
                             double* moment3, const uint32_t* sorted_index,
                             const uint32_t* cstart, const uint32_t* cend, uint32_t np,
                             uint32_t cRun);
/* cu_cube_hertzian_contact_force — cu_dem_def.cu:109-138, kernel
 * cu_dem_impl.cuh:329-378: the six planes (device array) in one pass, face
 * contacts only.  is_line_contact is unused by the reference kernel. */
int32_t xdem_b_cube_hertzian_contact_force(const xdem_b_plane_info* planes6, double kn,
                                           double vn, double ks, double vs, double mu,
                                           bool* is_line_contact, const double* pos4,
                                           const double* vel4, const double* omega4,
                                           double* force3, double* moment3, uint32_t np);
/* cu_shape_hertzian_contact_force — cu_dem_def.cu:140-232, kernels
 * cu_dem_impl.cuh:380-762 and reduce6 :820-847: force and moment are cleared (the reference's
 * per-step reset of its caller), then each particle takes its first face
 * contact with a triangle of the vertices in its 27 cells, else its first
 * edge/vertex contact; returns the sum of the particle forces (reduce6 block
 * tree, blocks summed in order) in out_body_force. */
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
                                            const uint32_t* cell_end, double out_body_force[3]);
/* cu_update_polygons_position — cu_dem_def.cu:234-238, kernel
 * cu_dem_impl.cuh:764-780: pos[i] = mpos + A local[i] (A row-major 3x3,
 * device memory). */
int32_t xdem_b_update_polygons_position(const double* mpos3, const double* A9,
                                        const double* local_vertice3, uint32_t nvertex,
                                        double* pos3);
/* cu_update_polygons_information — cu_dem_def.cu:240-244, kernel
 * cu_dem_impl.cuh:798-818: P, Q, R from the vertex indices, V = Q - P,
 * W = R - P, N = V x W. */
int32_t xdem_b_update_polygons_information(xdem_b_polygon* polygons, const uint32_t* indice3,
                                           const double* pos3, uint32_t ni);

#ifdef __cplusplus
}
#endif
#endif
