"""Palette JSON (LoadPalette / SavePalette, mixer/src/Serialization.cxx:26-53)
and getThinningMedium (mixer/src/Palette.cxx:453-463): host code of the
library, no device needed.  Fixtures: the reference's palette files
(data/paint_palettes/*.json, copied to tests/golden/palettes/)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PALETTES = os.path.join(HERE, "golden", "palettes")


@pytest.fixture(scope="module")
def hip():
    """The library's Python view, without a device context."""
    from painty_amd import _lib
    from painty_amd import hip as H
    _lib.load()
    return H


@pytest.mark.parametrize("name", ["lindemeier-measured", "curtis-watercolor", "thinning_medium"])
def test_reference_palettes_load_exactly(hip, name):
    text = open(os.path.join(PALETTES, name + ".json")).read()
    K, S = hip.palette_from_json(text)
    ref = json.loads(text)
    assert K.shape == (len(ref), 3)
    np.testing.assert_array_equal(K, [e["K"] for e in ref])
    np.testing.assert_array_equal(S, [e["S"] for e in ref])


@pytest.mark.parametrize("name", ["lindemeier-measured", "curtis-watercolor", "thinning_medium"])
def test_save_is_compact_json_and_round_trips(hip, name):
    K, S = hip.palette_from_json(open(os.path.join(PALETTES, name + ".json")).read())
    text = hip.palette_to_json(K, S)
    # the compact form with shortest round-trip decimals (Python's float repr)
    assert text == json.dumps([{"K": list(k), "S": list(s)} for k, s in zip(K, S)],
                              separators=(",", ":"))
    K2, S2 = hip.palette_from_json(text)
    np.testing.assert_array_equal(K2, K)
    np.testing.assert_array_equal(S2, S)


def test_save_formats_extreme_values(hip):
    rng = np.random.default_rng(1)
    K = np.concatenate([rng.uniform(0, 1, (20, 3)), [[1e-5, 1e16, 123456789.0],
                                                    [5e-324, 1.7976931348623157e308, -0.0]]])
    S = np.concatenate([rng.lognormal(0, 5, (20, 3)), [[1.0, 0.1, 1e-4], [2.0, 3e20, 7e-300]]])
    text = hip.palette_to_json(K, S)
    assert text == json.dumps([{"K": list(k), "S": list(s)} for k, s in zip(K, S)],
                              separators=(",", ":"))
    K2, S2 = hip.palette_from_json(text)
    np.testing.assert_array_equal(K2, K)
    np.testing.assert_array_equal(S2, S)


def test_load_accepts_whitespace_key_order_and_extra_keys(hip):
    text = ' [ {"S" : [1, 2.5e-1, 3], "name": {"a": [1, "x", null]}, "K":[0.5,0.25,1e0]} ]\n'
    K, S = hip.palette_from_json(text)
    np.testing.assert_array_equal(K, [[0.5, 0.25, 1.0]])
    np.testing.assert_array_equal(S, [[1.0, 0.25, 3.0]])
    assert hip.palette_from_json("[]")[0].shape == (0, 3)


@pytest.mark.parametrize("bad", ['{"K":[1,2,3]}', '[{"K":[1,2,3]}]', '[{"K":[1,2],"S":[1,2,3]}]',
                                 '[{"K":[1,2,3],"S":[1,2,3]}', '[{"K":[1,2,3],"S":[1,2,3]}] x',
                                 '[{"K":[1,2,3],"S":[a,2,3]}]'])
def test_load_rejects_malformed_palettes(hip, bad):
    with pytest.raises(hip.InvalidArgument):
        hip.palette_from_json(bad)


def test_thinning_medium_is_the_reference_data_file(hip):
    ref = json.load(open(os.path.join(PALETTES, "thinning_medium.json")))[0]
    K, S = hip.thinning_medium()
    np.testing.assert_array_equal(K, ref["K"])
    np.testing.assert_array_equal(S, ref["S"])
