"""The differentiable DNS step's element and side geometry
(semtex_amd.adjoint._geometry, host code) against the oracle's
Element::mapping restatement and side geometry on distorted quads."""
import numpy as np

from semtex_amd.mesh import rect_mesh


def test_geometry_matches_oracle_on_distorted_elements(oracle):
    from semtex_amd.adjoint import _geometry
    m = rect_mesh(3, 2, 6, 0.0, 2.0, -1.0, 1.0, bc=("D", "N", "D", "N"))
    # smooth distortion that keeps every Jacobian positive
    x, y = m.x.copy(), m.y.copy()
    m.x = x + 0.05 * np.sin(np.pi * y) * np.sin(np.pi * x / 2)
    m.y = y + 0.04 * np.sin(np.pi * x)
    g = _geometry(m)
    for e in range(m.nel):
        geo = oracle.mapping(m.np, m.x[e], m.y[e])[1]
        for key, q in (("rx", oracle.DRDX), ("ry", oracle.DRDY), ("sx", oracle.DSDX),
                       ("sy", oracle.DSDY), ("q4", oracle.Q4)):
            assert np.allclose(g[key][e].ravel(), geo[q], rtol=1e-12, atol=1e-12), (e, key)
    L = oracle.lib()
    for b in range(m.nbound):
        e, s = int(m.bnd_elmt[b]), int(m.bnd_side[b])
        geo = oracle.mapping(m.np, m.x[e], m.y[e])[1]
        area, nx, ny = np.zeros(m.np), np.zeros(m.np), np.zeros(m.np)
        L.or_side_geom(m.np, np.ascontiguousarray(geo.ravel()), np.ascontiguousarray(m.x[e]),
                       np.ascontiguousarray(m.y[e]), s, area, nx, ny)
        sl = slice(b * m.np, (b + 1) * m.np)
        assert np.allclose(g["barea"][sl], area, rtol=1e-12, atol=1e-14), b
        assert np.allclose(g["bnx"][sl], nx, atol=1e-13), b
        assert np.allclose(g["bny"][sl], ny, atol=1e-13), b
