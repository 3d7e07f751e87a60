"""ComputeScatteringAndAbsorption (KubelkaMunk.hxx:84-124) on the device in
FP64 against the FP64 oracle, and the round trip through ComputeReflectance
that defines it (a unit layer over black gives Rb, over white Rw)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _colours(n, seed):
    rng = np.random.default_rng(seed)
    Rb = rng.uniform(0.01, 0.6, (n, 3))
    Rw = Rb + rng.uniform(0.01, 0.39, (n, 3)) * (0.99 - Rb) / 0.4
    return Rb, np.minimum(Rw, 0.99)


def test_scattering_absorption_matches_oracle(hip, oracle):
    Rb, Rw = _colours(5000, 3)
    K, S = hip.compute_scattering_absorption(Rb, Rw)
    for i in range(0, 5000, 37):
        k, s = oracle.compute_scattering_absorption(Rb[i], Rw[i])
        np.testing.assert_allclose(K[i], k, rtol=1e-12)
        np.testing.assert_allclose(S[i], s, rtol=1e-12)


def test_scattering_absorption_round_trip(hip, oracle):
    Rb, Rw = _colours(200, 4)
    K, S = hip.compute_scattering_absorption(Rb, Rw)
    for i in range(200):
        np.testing.assert_allclose(oracle.compute_reflectance(K[i], S[i], [0, 0, 0], 1.0), Rb[i],
                                   rtol=1e-8)
        np.testing.assert_allclose(oracle.compute_reflectance(K[i], S[i], [1, 1, 1], 1.0), Rw[i],
                                   rtol=1e-8)


@pytest.mark.parametrize("rb,rw", [(0.7, 0.6), (0.0, 0.5), (0.2, 1.0), (0.3, 0.3), (-0.1, 0.5)])
def test_scattering_absorption_rejects_invalid_colours(hip, rb, rw):
    Rb, Rw = _colours(10, 5)
    Rb[4, 1], Rw[4, 1] = rb, rw
    with pytest.raises(hip.InvalidArgument):
        hip.compute_scattering_absorption(Rb, Rw)
