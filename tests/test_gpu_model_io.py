"""A model read from .xdm text runs on the FP64 GPU path and matches the
oracle step for step: cube-generator particles in a cube-object box, a ramp
plane under the lowest layer and an STL fin between two lattice columns (both
in contact from the start)."""
import numpy as np
import pytest

from metakonga_amd import model_io as mio
from metakonga_amd.xdem import Simulation
from test_model_io import XDM

pytestmark = pytest.mark.gpu

FIN = """solid fin
facet normal 0 0 1
 outer loop
  vertex 43.5 14 43.5
  vertex 63.5 14 43.5
  vertex 43.5 34 43.5
 endloop
endfacet
facet normal 1 0 0
 outer loop
  vertex 43.5 14 43.5
  vertex 43.5 34 43.5
  vertex 43.5 14 63.5
 endloop
endfacet
endsolid fin
"""


def test_xdm_model_steps_match_oracle(lib, orc, tmp_path):
    (tmp_path / "fin mesh.stl").write_text(FIN)
    text = XDM.replace("DIMENSION 0.05 0.05 0.05", "DIMENSION 0.08 0.08 0.08")
    for a, b in (("POINT_A 0 0.02 0", "POINT_A 0 0.0115 0"),
                 ("POINT_B 0.1 0.02 0", "POINT_B 0.1 0.0115 0"),
                 ("POINT_C 0.1 0.01 0.1", "POINT_C 0.1 0.0105 0.1"),
                 ("POINT_D 0 0.01 0.1", "POINT_D 0 0.0105 0.1")):
        text = text.replace(a, b)
    (tmp_path / "m.xdm").write_text(text)
    m = mio.read_xdm(tmp_path / "m.xdm")
    sc = m.to_scene(dt=1e-5, nsteps=30, grid_size=(32, 32, 32),
                    world_origin=(-0.05, -0.05, -0.05), seed=3)
    rng = np.random.default_rng(8)
    vel = rng.normal(size=(sc.np, 3)) * 0.5
    sim = Simulation(sc, lib)
    mass, iner = sim.mass_inertia()
    w = orc.World(sc, mass=mass, iner=iner)
    w.vel[:] = vel
    sim.set_state(vel=vel)
    for k in range(3):
        sim.step(10)
        w.step(10)
        st = sim.state()
        assert np.abs(st["pos"] - w.pos).max() / np.abs(w.pos[:, :3]).max() < 1e-13, k
        assert np.abs(st["vel"] - w.vel).max() / np.abs(w.vel).max() < 1e-9, k
        of = np.array([w.dpmi[0].force.x, w.dpmi[0].force.y, w.dpmi[0].force.z])
        g = sim.objects()[0].force
        assert np.abs(of).max() > 0
        assert np.abs(np.array([g.x, g.y, g.z]) - of).max() <= 1e-9 * np.abs(of).max(), k
    sim.close()
