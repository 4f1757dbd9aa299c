"""The drop-in boundary exercised the way the reference's own code calls it.

* The grid object (grid_base / neighborhood_cell, include/grid_base.h:11-46,
  source/neighborhood_cell.cpp:91-107) through its C-ABI entry points:
  create, set origin/cell size/grid size, initialize, detection with and
  without polygon spheres, then sortedID/cellStart/cellEnd — bit-exact
  against the oracle's hash + stable sort + cell ranges, over repeated
  detections on the same object.
* A caller compiled against include/xdem_compat.hpp only
  (tests/compat/compat_step.cpp: the reference's names, the order of
  dem_simulation::oneStepAnalysis and contactManager::deviceCollision, and
  the dpmi read-back of contact_particles_polygonObjects.cpp:469-477) runs
  whole steps of particles, a plane and two moving triangle objects.  Against
  the oracle composed in the same order: grid bit-exact, forces and the
  per-object reactions handed back in dpmi within 1e-12, positions after the
  steps within 1e-12 m.
"""
import ctypes as C
import os

import numpy as np
import pytest

from metakonga_amd import abi, scenes
from test_gpu_f64 import MAT, _grid, _triangle_scene, make_params
from xdem_testutil import rel_error

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
COMPAT_LIB = os.path.join(HERE, "compat", "libcompat_step.so")
TOL = 1e-12
G = 9.80665


def _read_u32(lib, ptr, n):
    out = np.empty(n, np.uint32)
    if n:
        abi.check(lib, lib.xdem_memcpy_dtoh(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), 4 * n))
    return out


@pytest.mark.parametrize("with_spheres", [False, True])
def test_grid_object_bitexact(lib, orc, with_spheres):
    pos, vel, om, dpi, sph, objs, rmax = _triangle_scene(seed=5)
    n = pos.shape[0]
    snp = sph.shape[0] if with_spheres else 0
    cs = 2 * rmax
    wo, gs = (-0.5, -0.5, -0.5), (64, 32, 64)
    p = make_params(n, wo, cs, gs)
    g = C.c_void_p()
    abi.check(lib, lib.xdem_grid_create(C.byref(g)))
    try:
        abi.check(lib, lib.xdem_grid_set_world_origin(g, *wo))
        abi.check(lib, lib.xdem_grid_set_cell_size(g, cs))
        abi.check(lib, lib.xdem_grid_set_grid_size(g, *gs))
        abi.check(lib, lib.xdem_grid_initialize(g, n + snp))
        ncell = lib.xdem_grid_ncell(g)
        assert ncell == gs[0] * gs[1] * gs[2]
        from xdem_testutil import DBuf
        dsph = DBuf(lib, sph)
        for shift in (0.0, 0.37 * cs, -1.4 * cs):
            q = pos.copy()
            q[:, :3] += shift
            dpos = DBuf(lib, q)
            abi.check(lib, lib.xdem_grid_detection(g, dpos.ptr, dsph.ptr if snp else None, n, snp))
            h, idx, cst, ce, srt = _grid(lib, orc, p, q, extra_sph=sph if snp else None)
            assert np.array_equal(_read_u32(lib, lib.xdem_grid_sorted_id(g), n + snp), srt)
            assert np.array_equal(_read_u32(lib, lib.xdem_grid_cell_start(g), ncell), cst)
            assert np.array_equal(_read_u32(lib, lib.xdem_grid_cell_end(g), ncell), ce)
            assert np.array_equal(_read_u32(lib, lib.xdem_grid_cell_id(g), n + snp), h)
        # more entries than initialize() reserved: refused, nothing written
        assert lib.xdem_grid_detection(g, dpos.ptr, dsph.ptr, n, snp + 1) != 0
    finally:
        lib.xdem_grid_destroy(g)


def _compat():
    assert os.path.exists(COMPAT_LIB), "tests/compat/libcompat_step.so not built"
    L = C.CDLL(COMPAT_LIB)
    vp = C.c_void_p
    u32 = C.c_uint32
    # prm, nsteps, 9 state arrays + cp_pp, planes, nplane, cp_wall, tris, sph4,
    # ntri, dpmi, nobj, cp_obj, sorted/cstart/cend out
    L.compat_steps.argtypes = ([vp, u32] + [vp] * 10 + [vp, u32, vp, vp, vp, u32, vp, u32, vp]
                               + [vp] * 3)
    L.compat_steps.restype = C.c_int
    return L


def _dpmi_rows(arr):
    """(force, moment) of each device_polygon_mass_info."""
    return (np.array([[o.force.x, o.force.y, o.force.z] for o in arr]),
            np.array([[o.moment.x, o.moment.y, o.moment.z] for o in arr]))


@pytest.mark.parametrize("nsteps", [1, 3])
def test_compat_header_steps(lib, orc, nsteps):
    pos, vel, om, dpi, sph, objs, rmax = _triangle_scene()
    n, nt = pos.shape[0], len(dpi)
    cs = 2 * rmax
    p = make_params(n, (-0.5, -0.5, -0.5), cs, (64, 64, 64))
    p.nsphere = nt
    m, I = scenes.mass_inertia(pos[:, 3], MAT.density)
    cp = scenes.contact_property(MAT, MAT, scenes.ContactSpec())
    cpw = scenes.contact_property(MAT, MAT, scenes.ContactSpec(friction=0.4))
    cps = [scenes.contact_property(MAT, MAT, scenes.ContactSpec(friction=0.3 + 0.1 * o))
           for o in range(2)]
    floor = scenes.plane4((-1, -0.006, -1), (1, -0.006, -1), (1, -0.006, 1), (-1, -0.006, 1))
    planes = abi.struct_array(abi.PlaneInfo, [floor])
    f0 = np.ascontiguousarray(m[:, None] * np.array([0.0, -G, 0.0])[None, :])

    # the compiled compat caller
    st = {k: np.ascontiguousarray(v.copy()) for k, v in dict(
        pos=pos, vel=vel, om=om, acc=np.zeros((n, 3)), al=np.zeros((n, 3)), f=f0,
        t=np.zeros((n, 3))).items()}
    dpmi = abi.struct_array(abi.PolygonMassInfo, objs)
    dpi_arr = abi.struct_array(abi.PolygonInfo, dpi)
    cpo = abi.struct_array(abi.ContactProperty, cps)
    srt_g = np.zeros(n + nt, np.uint32)
    cst_g = np.zeros(p.ncell, np.uint32)
    ce_g = np.zeros(p.ncell, np.uint32)
    P = abi.ptr
    rc = _compat().compat_steps(
        C.cast(C.pointer(p), C.c_void_p), nsteps, P(st["pos"]), P(st["vel"]), P(st["om"]),
        P(st["acc"]), P(st["al"]), P(st["f"]), P(st["t"]), P(m), P(I),
        C.cast(C.pointer(cp), C.c_void_p), C.cast(planes, C.c_void_p), 1,
        C.cast(C.pointer(cpw), C.c_void_p), C.cast(dpi_arr, C.c_void_p), P(sph), nt,
        C.cast(dpmi, C.c_void_p), 2, C.cast(cpo, C.c_void_p), P(srt_g), P(cst_g), P(ce_g))
    assert rc == 0

    # the oracle, composed in the same order
    r = {k: np.ascontiguousarray(v.copy()) for k, v in dict(
        pos=pos, vel=vel, om=om, acc=np.zeros((n, 3)), al=np.zeros((n, 3)), f=f0,
        t=np.zeros((n, 3))).items()}
    dpmi_ref = abi.struct_array(abi.PolygonMassInfo, objs)
    for k in range(nsteps):
        orc.vv_position(p, r["pos"], r["vel"], r["acc"])
        h, idx, cst, ce, srt = _grid(lib, orc, p, r["pos"], extra_sph=sph)
        with orc.force_magnitudes(n) as fm:
            orc.p2p(p, 1, r["pos"], r["vel"], r["om"], r["f"], r["t"], m, srt, cst, ce, cp)
            orc.plane(p, 1, floor, r["pos"], r["vel"], r["om"], r["f"], r["t"], m, cpw)
            orc.polygon(p, 1, dpi_arr, dpmi_ref, r["pos"], r["vel"], r["om"], r["f"], r["t"], m,
                        srt, cst, ce, cpo)
        f_last, t_last = r["f"].copy(), r["t"].copy()
        orc.vv_velocity(p, r["vel"], r["acc"], r["om"], r["al"], r["f"], r["t"], m, I)

    assert np.array_equal(srt_g, srt)
    assert np.array_equal(cst_g, cst)
    assert np.array_equal(ce_g, ce)
    scale = fm + m * G
    ferr = rel_error(st["f"], f_last, scale)
    terr = rel_error(st["t"], t_last, scale * pos[:, 3])
    print(f"compat caller, {nsteps} steps: force err {ferr:.2e}, torque err {terr:.2e}")
    assert ferr < TOL and terr < TOL
    assert np.count_nonzero(fm) > n // 2
    # the reactions the reference's caller hands to its bodies
    rf, rm = _dpmi_rows(dpmi)
    rf_ref, rm_ref = _dpmi_rows(dpmi_ref)
    fsum = fm.sum()
    assert np.abs(rf_ref).max() > 1e-6 * fsum
    assert np.abs(rf - rf_ref).max() <= TOL * fsum, (rf, rf_ref)
    assert np.abs(rm - rm_ref).max() <= TOL * fsum * 0.2, (rm, rm_ref)
    assert np.abs(st["pos"] - r["pos"]).max() < 1e-12
