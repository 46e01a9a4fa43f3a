"""Read-out and on-disk format: after the rhoU<CELL, true> read-out
(flb_store_moments) the GPU block written by flb_write_vtm gives the same
bytes as FreeLB's own vtmo::vtmWriter::WriteBinary for the same cavity
(fixtures written by oracle/_ref/vtk_ref, tests/golden/vtk/)."""
import os
import tempfile

import pytest

from freelb_amd import write_vtm
from gpu_case import gpu_lattice
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtk")


@pytest.mark.parametrize("lat,dims,steps,strip,name", [("d3q19", (8, 7, 6), 40, 1, "cav3"),
                                                        ("d2q9", (16, 13, 0), 100, 0, "cav2")])
def test_vtk_output_matches_freelb_bytes(lat, dims, steps, strip, name):
    case, p = O.cavity_case("hw", *dims)
    g = gpu_lattice(case, "f32", p)
    try:
        g.step(steps)
        g.store_moments(0xFF & ~O.VOID)
        mc = (-0.5, -0.5, -0.5 if lat == "d3q19" else 0.0)
        with tempfile.TemporaryDirectory() as tmp:
            write_vtm([g], [mc], [1.0], tmp, name, step=steps, strips=[strip], rho_name="Rho",
                      u_name="Velocity", flag_name="flag")
            for fn in (f"{name}{steps}.vtm", f"{name}_T{steps}_B0.vti"):
                mine = open(os.path.join(tmp, fn), "rb").read()
                ref = open(os.path.join(GOLDEN, fn), "rb").read()
                assert mine == ref, fn
    finally:
        g.close()
