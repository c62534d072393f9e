"""Host side of the core program (C1): the analytic C60 of generate_c60
(C1:228-252) against the reference build's output, and its geometry.

No GPU needed: fmd_generate_c60 is host code in libfmd.so."""
import numpy as np

from conftest import golden
from fuller_md_en_amd import system


def test_generate_c60_equals_reference():
    g = golden("c60_analytic.npz")
    m = system.c60_analytic()
    assert np.array_equal(m.coords, g["coords"])  # same vertices in the same order
    assert m.I0 == float(g["I0"]) and m.mass == float(g["Mmol"]) and m.rmol == float(g["Rmol"])


def test_generate_c60_is_a_truncated_icosahedron():
    """60 vertices on one sphere; every vertex has exactly three neighbours at
    the edge length 2 x 0.72 = 1.44 A (90 edges); I0 = (2/3) M R^2 for a shell."""
    m = system.c60_analytic()
    x = m.coords
    assert abs(x.mean(axis=0)).max() < 1e-14
    r = np.linalg.norm(x, axis=1)
    assert np.ptp(r) < 1e-13
    d = np.linalg.norm(x[:, None] - x[None], axis=2)
    edges = np.isclose(d, 1.44, atol=1e-12)
    assert edges.sum() == 180 and (edges.sum(axis=1) == 3).all()
    assert (d[~np.eye(60, dtype=bool)] > 1.44 - 1e-12).all()
    assert abs(m.I0 - 2.0 / 3.0 * m.mass * r[0] ** 2) < 1e-9
    assert abs(m.dmol - 2 * r[0]) < 1e-12  # antipodal vertices
