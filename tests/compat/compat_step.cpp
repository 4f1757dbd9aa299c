// compat_step.cpp — test infrastructure: a caller of the DEM step written only
// against include/xdem_compat.hpp (the reference's own names and objects), in
// the order the reference's GPU loop calls them:
//   dem_simulation::oneStepAnalysis   source/simulation/dem_simulation.cpp:447-469
//   contactManager::deviceCollision   source/contactManager.cpp:156-185
//   contact_particles_polygonObjects::cuda_collision
//                                     source/contact_particles_polygonObjects.cpp:469-477
//                                     (dpmi copied back after every call)
// Host arrays in, host arrays out; device memory through the C ABI.
#include <cstdint>
#include <cstring>
#include <vector>

#include "xdem_compat.hpp"

namespace {
template <typename T>
T* upload(const T* h, size_t n) {
    void* d = nullptr;
    xdem_compat::check(xdem_malloc(&d, sizeof(T) * (n ? n : 1)), "malloc");
    if (n) xdem_compat::check(xdem_memcpy_htod(d, h, sizeof(T) * n), "htod");
    return static_cast<T*>(d);
}
template <typename T>
void download(T* h, const T* d, size_t n) {
    if (n) xdem_compat::check(xdem_memcpy_dtoh(h, d, sizeof(T) * n), "dtoh");
}
}  // namespace

extern "C" int compat_steps(const device_parameters* prm_in, uint32_t nsteps, double* pos4,
                            double* vel3, double* omega3, double* acc3, double* alpha3,
                            double* force3, double* moment3, const double* mass,
                            const double* iner, const device_contact_property* cp_pp,
                            const device_plane_info* planes, uint32_t nplane,
                            const device_contact_property* cp_wall,
                            const device_polygon_info* tris, const double* sph4, uint32_t ntri,
                            device_polygon_mass_info* dpmi, uint32_t nobj,
                            const device_contact_property* cp_obj, uint32_t* sorted_out,
                            uint32_t* cstart_out, uint32_t* cend_out) {
    device_parameters prm = *prm_in;
    const unsigned int np = prm.np;
    prm.nsphere = ntri;
    setSymbolicParameter(&prm);
    double* dpos = upload(pos4, 4ull * np);
    double* dvel = upload(vel3, 3ull * np);
    double* domega = upload(omega3, 3ull * np);
    double* dacc = upload(acc3, 3ull * np);
    double* dalpha = upload(alpha3, 3ull * np);
    double* dforce = upload(force3, 3ull * np);
    double* dmoment = upload(moment3, 3ull * np);
    double* dmass = upload(mass, np);
    double* diner = upload(iner, np);
    device_contact_property* dcp = upload(cp_pp, 1);
    device_contact_property* dcpw = upload(cp_wall, 1);
    device_plane_info* dplanes = upload(planes, nplane);
    device_polygon_info* dpi = upload(tris, ntri);
    double* dsph = upload(sph4, 4ull * ntri);
    device_polygon_mass_info* ddpmi = upload(dpmi, nobj);
    device_contact_property* dcpo = upload(cp_obj, nobj);

    neighborhood_cell dtor;
    dtor.setWorldOrigin(prm.world_origin.x, prm.world_origin.y, prm.world_origin.z);
    dtor.setCellSize(prm.cell_size);
    dtor.setGridSize(prm.grid_size.x, prm.grid_size.y, prm.grid_size.z);
    dtor.initialize(np + ntri);
    velocity_verlet itor;
    for (uint32_t k = 0; k < nsteps; ++k) {
        itor.updatePosition(dpos, dvel, dacc, np);
        dtor.detection(dpos, ntri ? dsph : nullptr, np, ntri);
        cu_calculate_p2p(1, dpos, dvel, domega, dforce, dmoment, dmass, dtor.sortedID(),
                         dtor.cellStart(), dtor.cellEnd(), dcp, np);
        for (uint32_t q = 0; q < nplane; ++q)
            cu_plane_contact_force(1, dplanes + q, dpos, dvel, domega, dforce, dmoment, dmass, np,
                                   dcpw);
        if (ntri) {
            cu_particle_polygonObject_collision(1, dpi, dsph, ddpmi, dpos, dvel, domega, dforce,
                                                dmoment, dmass, dtor.sortedID(), dtor.cellStart(),
                                                dtor.cellEnd(), dcpo, np);
            download(dpmi, ddpmi, nobj);  // handed to the bodies by the reference caller
        }
        if (k + 1 == nsteps) {
            // the step's forces, before the Verlet update resets them
            download(force3, dforce, 3ull * np);
            download(moment3, dmoment, 3ull * np);
            download(sorted_out, dtor.sortedID(), np + ntri);
            download(cstart_out, dtor.cellStart(), dtor.nCell());
            download(cend_out, dtor.cellEnd(), dtor.nCell());
        }
        itor.updateVelocity(dvel, dacc, domega, dalpha, dforce, dmoment, dmass, diner, np);
    }
    download(pos4, dpos, 4ull * np);
    download(vel3, dvel, 3ull * np);
    download(omega3, domega, 3ull * np);
    download(acc3, dacc, 3ull * np);
    download(alpha3, dalpha, 3ull * np);
    for (void* p : {(void*)dpos, (void*)dvel, (void*)domega, (void*)dacc, (void*)dalpha,
                    (void*)dforce, (void*)dmoment, (void*)dmass, (void*)diner, (void*)dcp,
                    (void*)dcpw, (void*)dplanes, (void*)dpi, (void*)dsph, (void*)ddpmi,
                    (void*)dcpo})
        xdem_free(p);
    return 0;
}
