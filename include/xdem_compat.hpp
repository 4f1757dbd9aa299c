// xdem_compat.hpp — the reference's host C++ API over the xdem C ABI.
//
// A drop-in for include/mphysics_cuda_dec.cuh (Set A, FP64) and for the
// grid_base / neighborhood_cell / velocity_verlet objects
// (include/grid_base.h, include/neighborhood_cell.h, include/dem_integrator.h):
// same names, argument order and meaning.  A failing call prints
// xdem_last_error() and aborts, as the reference's checkCudaErrors exits.
//
// Include this instead of mphysics_cuda_dec.cuh in a translation unit that
// does not include CUDA/HIP headers (the reference's double3/double4/uint3
// names are provided here; define XDEM_COMPAT_NO_VECTOR_TYPES to skip them).
#pragma once

#include <cstdio>
#include <cstdlib>

#include "xdem.h"

#ifndef XDEM_COMPAT_NO_VECTOR_TYPES
using double3 = xdem_double3;
using double4 = xdem_double4;
using uint3 = xdem_uint3;
#endif
using device_parameters = xdem_device_parameters;
using device_polygon_info = xdem_device_polygon_info;
using device_plane_info = xdem_device_plane_info;
using device_polygon_mass_info = xdem_device_polygon_mass_info;
using device_cylinder_info = xdem_device_cylinder_info;
using device_contact_property = xdem_device_contact_property;
using device_force_constant = xdem_device_force_constant;

namespace xdem_compat {
inline void check(int32_t rc, const char* what) {
    if (rc != XDEM_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, xdem_last_error());
        std::abort();
    }
}
}  // namespace xdem_compat

// ------------------------------------- mphysics_cuda_dec.cuh:110-161 (FP64)
inline void setSymbolicParameter(device_parameters* h_paras) {
    xdem_compat::check(xdem_set_symbolic_parameter(h_paras), "setSymbolicParameter");
}
inline void vv_update_position(double* pos, double* vel, double* acc, unsigned int np) {
    xdem_compat::check(xdem_vv_update_position_f64(pos, vel, acc, np), "vv_update_position");
}
inline void vv_update_velocity(double* vel, double* acc, double* omega, double* alpha,
                               double* force, double* moment, double* mass, double* iner,
                               unsigned int np) {
    xdem_compat::check(
        xdem_vv_update_velocity_f64(vel, acc, omega, alpha, force, moment, mass, iner, np),
        "vv_update_velocity");
}
inline void cu_calculateHashAndIndex(unsigned int* hash, unsigned int* index, double* pos,
                                     unsigned int np) {
    xdem_compat::check(xdem_calculate_hash_and_index_f64(hash, index, pos, np),
                       "cu_calculateHashAndIndex");
}
inline void cu_calculateHashAndIndexForPolygonSphere(unsigned int* hash, unsigned int* index,
                                                     unsigned int sid, unsigned int nsphere,
                                                     double* sphere) {
    xdem_compat::check(
        xdem_calculate_hash_and_index_polygon_sphere_f64(hash, index, sid, nsphere, sphere),
        "cu_calculateHashAndIndexForPolygonSphere");
}
inline void cu_reorderDataAndFindCellStart(unsigned int* hash, unsigned int* index,
                                           unsigned int* cstart, unsigned int* cend,
                                           unsigned int* sorted_index, unsigned int np,
                                           unsigned int ncell) {
    xdem_compat::check(
        xdem_reorder_data_and_find_cell_start(hash, index, cstart, cend, sorted_index, np, ncell),
        "cu_reorderDataAndFindCellStart");
}
inline void cu_calculate_p2p(const int tcm, double* pos, double* vel, double* omega, double* force,
                             double* moment, double* mass, unsigned int* sorted_index,
                             unsigned int* cstart, unsigned int* cend, device_contact_property* cp,
                             unsigned int np) {
    xdem_compat::check(xdem_calculate_p2p_f64(tcm, pos, vel, omega, force, moment, mass,
                                              sorted_index, cstart, cend, cp, np),
                       "cu_calculate_p2p");
}
inline void cu_plane_contact_force(const int tcm, device_plane_info* plan, double* pos, double* vel,
                                   double* omega, double* force, double* moment, double* mass,
                                   unsigned int np, device_contact_property* cp) {
    xdem_compat::check(
        xdem_plane_contact_force_f64(tcm, plan, pos, vel, omega, force, moment, mass, np, cp),
        "cu_plane_contact_force");
}
inline void cu_cube_contact_force(const int tcm, device_plane_info* plan, double* pos, double* vel,
                                  double* omega, double* force, double* moment, double* mass,
                                  unsigned int np, device_contact_property* cp) {
    xdem_compat::check(
        xdem_cube_contact_force_f64(tcm, plan, pos, vel, omega, force, moment, mass, np, cp),
        "cu_cube_contact_force");
}
// The reference accumulates per-object reactions into dpmi with a data race
// and never zeroes them; here dpmi[].force/moment receive this call's ordered
// reaction sums (the objects are those the triangles name; the triangle count
// is the parameter block's nsphere), which is what the reference's caller
// copies back and hands to its bodies.
inline void cu_particle_polygonObject_collision(const int tcm, device_polygon_info* dpi,
                                                double* dsph, device_polygon_mass_info* dpmi,
                                                double* pos, double* vel, double* omega,
                                                double* force, double* moment, double* mass,
                                                unsigned int* sidx, unsigned int* cstart,
                                                unsigned int* cend, device_contact_property* cp,
                                                unsigned int np) {
    xdem_compat::check(xdem_particle_polygon_object_collision_f64(
                           tcm, dpi, dsph, dpmi, XDEM_NOBJECT_AUTO, 0, pos, vel, omega, force,
                           moment, mass, sidx,
                           cstart, cend, cp, np),
                       "cu_particle_polygonObject_collision");
}
inline void cu_cylinder_hertzian_contact_force(const int tcm, device_cylinder_info* cyl,
                                               double* pos, double* vel, double* omega,
                                               double* force, double* moment, double* mass,
                                               unsigned int np, device_contact_property* cp,
                                               double3* mpos, double3* mf, double3* mm,
                                               double3& _mf, double3& _mm) {
    xdem_compat::check(
        xdem_cylinder_hertzian_contact_force_f64(tcm, cyl, pos, vel, omega, force, moment, mass,
                                                 np, cp, mpos, mf, mm, &_mf, &_mm),
        "cu_cylinder_hertzian_contact_force");
}
inline double3 reductionD3(double3* in, unsigned int np) {
    double3 r{0, 0, 0};
    xdem_compat::check(xdem_reduction_d3(in, np, &r), "reductionD3");
    return r;
}
inline void cu_updatePolygonObjectData(device_polygon_mass_info* dpmi, double* vList,
                                       double* sphere, device_polygon_info* dpi,
                                       unsigned int ntriangle) {
    xdem_compat::check(xdem_update_polygon_object_data_f64(dpmi, vList, sphere, dpi, ntriangle),
                       "cu_updatePolygonObjectData");
}

// ------------------------- grid_base / neighborhood_cell (include/grid_base.h)
class neighborhood_cell {
public:
    neighborhood_cell() { xdem_compat::check(xdem_grid_create(&g_), "grid_create"); }
    ~neighborhood_cell() { xdem_grid_destroy(g_); }
    neighborhood_cell(const neighborhood_cell&) = delete;
    neighborhood_cell& operator=(const neighborhood_cell&) = delete;

    void setWorldOrigin(double x, double y, double z) {
        xdem_compat::check(xdem_grid_set_world_origin(g_, x, y, z), "setWorldOrigin");
    }
    void setCellSize(double cs) { xdem_compat::check(xdem_grid_set_cell_size(g_, cs), "setCellSize"); }
    void setGridSize(unsigned int x, unsigned int y, unsigned int z) {
        xdem_compat::check(xdem_grid_set_grid_size(g_, x, y, z), "setGridSize");
    }
    void initialize(unsigned int np) { xdem_compat::check(xdem_grid_initialize(g_, np), "initialize"); }
    // grid_base::detection (GPU branch of neighborhood_cell::detection)
    void detection(double* pos = nullptr, double* spos = nullptr, unsigned int np = 0,
                   unsigned int snp = 0) {
        xdem_compat::check(xdem_grid_detection(g_, pos, spos, np, snp), "detection");
    }
    unsigned int nCell() const { return xdem_grid_ncell(g_); }
    unsigned int* sortedID() { return xdem_grid_sorted_id(g_); }
    unsigned int* cellStart() { return xdem_grid_cell_start(g_); }
    unsigned int* cellEnd() { return xdem_grid_cell_end(g_); }
    void getCellNumber(double x, double y, double z, int out[3]) const {
        xdem_compat::check(xdem_grid_get_cell_number(g_, x, y, z, out), "getCellNumber");
    }
    unsigned int getHash(const int c3[3]) const { return xdem_grid_get_hash(g_, c3); }

private:
    xdem_grid* g_ = nullptr;
};

// ------------------------------ velocity_verlet (include/dem_integrator.h:16-28)
class velocity_verlet {
public:
    void updatePosition(double* dpos, double* dvel, double* dacc, unsigned int np) {
        vv_update_position(dpos, dvel, dacc, np);
    }
    void updateVelocity(double* dvel, double* dacc, double* domega, double* dalpha, double* dforce,
                        double* dmoment, double* dmass, double* dinertia, unsigned int np) {
        vv_update_velocity(dvel, dacc, domega, dalpha, dforce, dmoment, dmass, dinertia, np);
    }
};
