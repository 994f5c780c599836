"""Initial states of the benchmark cases (CPU)."""
import numpy as np

from hfem_amd.problem import make_problem


def test_taylor_green_state_is_the_closed_form():
    """Without noise the nodal state is the Taylor-Green vortex: velocity
    0.1 (sin X cos Y cos Z, -cos X sin Y cos Z, 0), unit density and the
    pressure 1/(gamma M^2) + 0.1^2/16 (cos 2X + cos 2Y)(cos 2Z + 2)."""
    pr = make_problem("hex", 3, 2, seed=0, state="taylor_green", amp=0.0)
    x = pr.mesh.dgnodes                                   # (ne, npv, 3)
    X, Y, Z = (2 * np.pi * x[..., i] for i in range(3))
    U = pr.UDG
    rho = U[:, 0]
    u, v, w = (U[:, 1 + i] / rho for i in range(3))
    g = 1.4
    p = (g - 1) * (U[:, 4] - 0.5 * rho * (u * u + v * v + w * w))
    assert np.allclose(rho, 1.0, atol=0, rtol=1e-15)
    assert np.allclose(u, 0.1 * np.sin(X) * np.cos(Y) * np.cos(Z), atol=1e-15)
    assert np.allclose(v, -0.1 * np.cos(X) * np.sin(Y) * np.cos(Z), atol=1e-15)
    assert np.abs(w).max() == 0.0
    p_ref = 1 / (g * 0.3 ** 2) + 0.01 / 16 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    assert np.allclose(p, p_ref, atol=1e-12, rtol=0)


def test_taylor_green_noise_is_bounded_and_seeded():
    a = make_problem("hex", 3, 2, seed=4, state="taylor_green", amp=1e-3)
    b = make_problem("hex", 3, 2, seed=4, state="taylor_green", amp=1e-3)
    c = make_problem("hex", 3, 2, seed=4, state="taylor_green", amp=0.0)
    assert np.array_equal(a.UDG, b.UDG)
    d = np.abs(a.UDG[:, :5] - c.UDG[:, :5])
    assert 0 < d.max() <= 1e-3
