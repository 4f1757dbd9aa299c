// xdem_compat_setb.hpp — the reference's older host C++ API over the xdem C ABI.
//
// A drop-in for include/cu_dem_dec.cuh (the parSIM generation's entry points):
// same names, argument order and meaning, forwarding to the xdem_b_* functions
// of xdem_setb.h.  A failing call prints xdem_last_error() and aborts, as the
// reference's checkCudaErrors exits.  Include it instead of cu_dem_dec.cuh (a
// translation unit includes either this header or xdem_compat.hpp: both
// sets use the same reference names for different structs).
#pragma once

#include <cstdio>
#include <cstdlib>

#include "xdem_setb.h"

#ifndef XDEM_COMPAT_NO_VECTOR_TYPES
using double3 = xdem_double3;
using double4 = xdem_double4;
using uint3 = xdem_uint3;
#endif
using device_parameters = xdem_b_device_parameters;
using device_plane_info = xdem_b_plane_info;
using cu_polygon = xdem_b_polygon;
using cu_mass_info = xdem_b_mass_info;

namespace xdem_compat_setb {
inline void check(int32_t rc, const char* what) {
    if (rc != XDEM_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, xdem_last_error());
        std::abort();
    }
}
}  // namespace xdem_compat_setb

// ------------------------------------------------------ cu_dem_dec.cuh:95-113
inline void setSymbolicParameter(device_parameters* h_paras) {
    xdem_compat_setb::check(xdem_b_set_symbolic_parameter(h_paras), "setSymbolicParameter");
}
inline void vv_update_position(double* pos, double* vel, double* acc, unsigned int np) {
    xdem_compat_setb::check(xdem_b_vv_update_position(pos, vel, acc, np), "vv_update_position");
}
inline void vv_update_velocity(double* vel, double* acc, double* omega, double* alpha,
                               double* force, double* moment, unsigned int np) {
    xdem_compat_setb::check(xdem_b_vv_update_velocity(vel, acc, omega, alpha, force, moment, np),
                            "vv_update_velocity");
}
inline void cu_calculateHashAndIndex(unsigned int* hash, unsigned int* index, double* pos,
                                     unsigned int np) {
    xdem_compat_setb::check(xdem_b_calculate_hash_and_index(hash, index, pos, np),
                            "cu_calculateHashAndIndex");
}
inline void cu_reorderDataAndFindCellStart(unsigned int* hash, unsigned int* index,
                                           unsigned int* cstart, unsigned int* cend,
                                           unsigned int* sorted_index, unsigned int np,
                                           unsigned int nc) {
    xdem_compat_setb::check(
        xdem_b_reorder_data_and_find_cell_start(hash, index, cstart, cend, sorted_index, np, nc),
        "cu_reorderDataAndFindCellStart");
}
inline void cu_calculate_p2p(double* pos, double* vel, double* acc, double* omega, double* alpha,
                             double* force, double* moment, unsigned int* sorted_index,
                             unsigned int* cstart, unsigned int* cend, unsigned int np,
                             unsigned int cRun = 0) {
    xdem_compat_setb::check(xdem_b_calculate_p2p(pos, vel, acc, omega, alpha, force, moment,
                                                 sorted_index, cstart, cend, np, cRun),
                            "cu_calculate_p2p");
}
inline void cu_cube_hertzian_contact_force(device_plane_info* plans, double kn, double vn,
                                           double ks, double vs, double mu, bool* isLineContact,
                                           double* pos, double* vel, double* omega, double* force,
                                           double* moment, unsigned int np) {
    xdem_compat_setb::check(xdem_b_cube_hertzian_contact_force(plans, kn, vn, ks, vs, mu,
                                                               isLineContact, pos, vel, omega,
                                                               force, moment, np),
                            "cu_cube_hertzian_contact_force");
}
inline double3 cu_shape_hertzian_contact_force(
    cu_mass_info* minfo, cu_polygon* polygons, unsigned int* id_set, unsigned int* poly_start,
    unsigned int* poly_end, double kn, double vn, double ks, double vs, double mu,
    bool* isLineContact, double* pos, double* vel, double* omega, double* force, double* moment,
    double3* bforce, unsigned int np, unsigned int* sorted_id, unsigned int* cell_start,
    unsigned int* cell_end) {
    double r[3];
    xdem_compat_setb::check(
        xdem_b_shape_hertzian_contact_force(minfo, polygons, id_set, poly_start, poly_end, kn, vn,
                                            ks, vs, mu, isLineContact, pos, vel, omega, force,
                                            moment, reinterpret_cast<double*>(bforce), np,
                                            sorted_id, cell_start, cell_end, r),
        "cu_shape_hertzian_contact_force");
    return double3{r[0], r[1], r[2]};
}
inline void cu_update_polygons_position(double3* mpos, double* A, double3* local_vertice,
                                        unsigned int nvertex, double3* pos) {
    xdem_compat_setb::check(
        xdem_b_update_polygons_position(reinterpret_cast<const double*>(mpos), A,
                                        reinterpret_cast<const double*>(local_vertice), nvertex,
                                        reinterpret_cast<double*>(pos)),
        "cu_update_polygons_position");
}
inline void cu_update_polygons_information(cu_polygon* polygons, uint3* indice, double3* pos,
                                           unsigned int ni) {
    xdem_compat_setb::check(
        xdem_b_update_polygons_information(polygons, reinterpret_cast<const uint32_t*>(indice),
                                           reinterpret_cast<const double*>(pos), ni),
        "cu_update_polygons_information");
}
inline void cu_mergedata(double* data, double* pos, unsigned int np, double3* vertice,
                         unsigned int snp) {
    xdem_compat_setb::check(
        xdem_b_mergedata(data, pos, np, reinterpret_cast<const double*>(vertice), snp),
        "cu_mergedata");
}
