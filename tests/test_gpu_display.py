"""Renderer::render (Renderer.hxx:60-156) and imSave's pixel transform
(io/src/ImageIO.cxx:158-183) on the GPU against the FP64 oracle: the lit
composite within 1e-5 relative per channel; the 8- and 16-bit encodings
bit for bit (the encode runs in FP64 on the float input)."""
import numpy as np
import pytest

from painty_amd import workload as W
from tests.parity import assert_close, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _canvas(hip, rows, cols, seed, vmax=0.5):
    rng = np.random.default_rng(seed)
    K = rng.uniform(0.01, 3.0, (rows, cols, 3)).astype(np.float32)
    S = rng.uniform(0.1, 3.0, (rows, cols, 3)).astype(np.float32)
    V = (rng.uniform(0.0, vmax, (rows, cols)) * (rng.uniform(size=(rows, cols)) > 0.3))
    V = V.astype(np.float32)
    R0 = rng.uniform(0.3, 1.0, (rows, cols, 3)).astype(np.float32)
    cv = hip.Canvas(rows, cols)
    cv.getPaintLayer().setBuffers(K, S, V)
    cv.setR0(R0)
    return cv, (K, S, V, R0)


@pytest.mark.parametrize("rows,cols,vmax", [(1, 1, 0.5), (1, 37, 0.5), (29, 1, 0.5),
                                            (64, 96, 0.5), (200, 150, 30.0), (2, 2, 5.0)])
def test_render_matches_oracle(hip, oracle, rows, cols, vmax):
    cv, (K, S, V, R0) = _canvas(hip, rows, cols, rows * 1000 + cols, vmax)
    got = hip.Renderer.render(cv)
    ref = oracle.render(K, S, V, R0)
    assert_close("render/%dx%d" % (rows, cols), got, ref)


def test_render_after_stamping(hip, oracle):
    """The lit composite of a stamped canvas (a bumpy height field)."""
    wl = W.small(96, 112, 40, 8, 5, pickup=True, noisy=True)
    ocv, _ = run_oracle(wl)
    gcv, _ = run_gpu(wl, mode=2)
    ref = oracle.render(ocv.K, ocv.S, ocv.V, ocv.R0)
    assert_close("render/stamped", hip.Renderer.render(gcv), ref)


@pytest.mark.parametrize("bits", [8, 16])
@pytest.mark.parametrize("to_srgb", [True, False])
def test_encode_bit_exact(hip, oracle, bits, to_srgb):
    rng = np.random.default_rng(bits + to_srgb)
    x = rng.uniform(-0.1, 1.1, (300, 200, 3)).astype(np.float32)
    scale = 65535.0 if bits == 16 else 255.0
    edge = np.array([0.0, 1.0, -0.0, 0.00313066844250063, 0.0031306685, 2.0, -1.0,
                     0.5 / scale, 1.5 / scale, 2.5 / scale, 127.5 / scale], np.float32)
    x[0, : len(edge) // 3 * 1, :] = edge[: len(edge) // 3 * 3].reshape(-1, 3)
    got = hip.encode(x, to_srgb=to_srgb, bits=bits)
    ref = oracle.encode(x.astype(np.float64), to_srgb=to_srgb, bits=bits)
    assert got.dtype == (np.uint16 if bits == 16 else np.uint8)
    np.testing.assert_array_equal(got, ref)
