"""Set-up helpers of the benchmark workloads (host side, no GPU): the divergence-free noise of
the Taylor-Green case and the Courant-number time step."""
import numpy as np

import prism_amd as P
from prism_amd import cases


def test_solenoidal_noise_is_divergence_free():
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 2 * np.pi, (200, 3))
    h = 1e-5

    def u(p):
        return cases.solenoidal_noise(p[:, 0], p[:, 1], p[:, 2], 1.0, seed=4, kmax=2)

    div = np.zeros(len(pts))
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        div += (u(pts + e)[:, i] - u(pts - e)[:, i]) / (2 * h)
    grad_scale = np.abs((u(pts + h) - u(pts - h)) / (2 * h)).max()
    assert grad_scale > 0.1
    assert np.abs(div).max() <= 1e-8 * grad_scale  # central-difference truncation ~h^2


def test_solenoidal_noise_rms_and_seed():
    """The scale is fixed by the coefficients (the box average), not by the sample points: on a
    uniform grid of the periodic box the RMS per component averages to sigma."""
    g = np.arange(16) * (2 * np.pi / 16)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    a = cases.solenoidal_noise(X.ravel(), Y.ravel(), Z.ravel(), 1e-3, seed=4)
    assert abs(np.sqrt((a ** 2).mean()) - 1e-3) <= 1e-12
    b = cases.solenoidal_noise(X.ravel(), Y.ravel(), Z.ravel(), 1e-3, seed=4)
    c = cases.solenoidal_noise(X.ravel(), Y.ravel(), Z.ravel(), 1e-3, seed=5)
    assert np.array_equal(a, b)
    assert not np.allclose(a, c)


def test_courant_dt_puts_the_maximum_at_the_target(ref):
    c = cases.taylor_green(ref, n=3, tets=True)
    co = P.courant_number(c["mdot"], c["dt"])
    assert abs(co.max() - 0.5) <= 1e-12
    # Co is linear in dt (courant.cpp:7-24)
    assert np.allclose(P.courant_number(c["mdot"], 2 * c["dt"]), 2 * co, rtol=1e-14)
