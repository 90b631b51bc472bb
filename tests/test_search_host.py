"""Host-side pieces of the quant-field search: the per-block perceptual
weights (CPU)."""
import numpy as np

from pik_amd.search import block_weights


def test_block_weights_follow_the_cube_root_slope():
    lin = np.zeros((3, 16, 24), np.float32)
    lin[:, :, :8] = 2.0     # dark blocks
    lin[:, :, 8:16] = 50.0
    lin[:, :, 16:] = 200.0  # bright blocks
    w = block_weights(lin)
    assert w.shape == (2, 3)
    assert abs(float(w.mean()) - 1.0) < 1e-6
    assert (w[:, 0] > w[:, 1]).all() and (w[:, 1] > w[:, 2]).all()
    # ratio of two columns = ratio of (Y + 8)^(-4/3)
    r = (10.0 / 58.0) ** (-4.0 / 3)
    assert abs(float(w[0, 0] / w[0, 1]) - r) < 1e-4 * r


def test_block_weights_pad_ragged_sizes():
    lin = np.full((3, 13, 7), 30.0, np.float32)
    w = block_weights(lin)
    assert w.shape == (2, 1)
    assert np.allclose(w, 1.0)


def test_search_field_fixture_is_the_references():
    """tests/golden/search_fields_1024.npz holds the reference's
    InitialQuantField of the seeded search images (checked on seed 0)."""
    import os
    import pytest
    from pik_amd import synth
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden",
                             "search_fields_1024.npz"))
    assert z["fields"].shape == (16, 128, 128)
    _, _, orig, _ = synth.make_search_inputs(
        1024, 1024, seed=int(z["seeds"][0]), num_rects=int(z["num_rects"]),
        target=float(z["target"]))
    want = ref.initial_quant_field(orig, float(z["target"]))
    assert np.array_equal(z["fields"][0], want)
    assert float(z["quant_dc"]) == ref.initial_quant_dc(float(z["target"]))
