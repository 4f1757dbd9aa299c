// xdem_f64.hpp — launchers of the plain FP64 kernels (xdem_f64.hip), shared by
// the Set A entry points and the host-side simulation object.
#pragma once
#include "xdem_common.hpp"

namespace xdem {

int32_t launch_vv_position(const KParams& k, double* pos, double* vel, double* acc, uint32_t np,
                           hipStream_t s);
int32_t launch_vv_velocity(const KParams& k, double* vel, double* acc, double* omega,
                           double* alpha, double* force, double* moment, const double* mass,
                           const double* iner, uint32_t np, hipStream_t s);
int32_t launch_hash(const KParams& k, uint32_t* hash, uint32_t* index, const double* pos,
                    uint32_t sid, uint32_t n, hipStream_t s);
int32_t launch_p2p(int tcm, const KParams& k, const double* pos, const double* vel,
                   const double* omega, double* force, double* moment, const double* mass,
                   const uint32_t* sorted, const uint32_t* cstart, const uint32_t* cend,
                   const xdem_device_contact_property* cp, uint32_t np, uint32_t nsorted,
                   hipStream_t s, float4* pfs = nullptr);
int32_t launch_planes(int tcm, const KParams& k, const xdem_device_plane_info* planes,
                      const xdem_device_contact_property* cps, uint32_t nplane, const double* pos,
                      const double* vel, const double* omega, double* force, double* moment,
                      const double* mass, uint32_t np, hipStream_t s);
int32_t launch_cylinder(int tcm, const KParams& k, const xdem_device_cylinder_info* cy,
                        const double* pos, const double* vel, const double* omega, double* force,
                        double* moment, const double* mass, uint32_t np,
                        const xdem_device_contact_property* cp, const xdem_double3* mpos,
                        xdem_double3* mf, xdem_double3* mm, hipStream_t s);
int32_t launch_drum(int tcm, const KParams& k, const xdem_drum_info* dr, const double* pos,
                    const double* vel, const double* omega, double* force, double* moment,
                    const double* mass, uint32_t np, const xdem_device_contact_property* cp,
                    xdem_double3* mf, xdem_double3* mm, hipStream_t s);
int32_t launch_put_plane(xdem_device_plane_info* dst, const xdem_device_plane_info& v,
                         hipStream_t s);
int32_t launch_put_object(xdem_device_polygon_mass_info* dst,
                          const xdem_device_polygon_mass_info& v, hipStream_t s);
int32_t launch_polygon(int tcm, const KParams& k, const xdem_device_polygon_info* dpi,
                       xdem_device_polygon_mass_info* dpmi, uint32_t nobj, const double* pos,
                       const double* vel, const double* omega, double* force, double* moment,
                       const double* mass, const uint32_t* sorted, const uint32_t* cstart,
                       const uint32_t* cend, const xdem_device_contact_property* cp, uint32_t np,
                       hipStream_t s);
int32_t launch_update_polygon(const xdem_device_polygon_mass_info* dpmi, const double* vList,
                              double* sphere, xdem_device_polygon_info* dpi, uint32_t ntri,
                              hipStream_t s);
int32_t reduction_d3(const d3* in, uint32_t n, d3* out, hipStream_t s);
int32_t launch_pairs(const KParams& k, const double* pos, const uint32_t* sorted,
                     const uint32_t* cstart, const uint32_t* cend, uint32_t np, uint32_t max_per,
                     uint32_t* counts, uint32_t* partners, hipStream_t s);

}  // namespace xdem
