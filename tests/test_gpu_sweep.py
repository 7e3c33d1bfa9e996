"""A reduced surrogate-guided sweep end to end on the GPU: batch-engine data,
training on the library's BF16 kernels, scoring, BD verification. The
trained network must beat the no-learning baseline (the training mean) on
the radial density, the quantity the parameters determine at these run
lengths."""
import numpy as np
import pytest

from gpu_common import step
from tfm_amd import sweep

pytestmark = pytest.mark.gpu


def test_reduced_sweep_learns_the_radial_density():
    step("reduced sweep: 320 sets x 3000 steps")
    out = sweep.run(["--no-tune", "--train", "320", "--warmup", "500", "--steps", "2500", "--interval", "100", "--epochs", "120",
                     "--candidates", "512", "--top", "32", "--single", "1", "--batch", "160"])
    err = out["contact_map_error"]
    step(f"  radial density MAE {err['radial_density_mae']}, map error {err['surrogate_validation']:.3f} "
         f"vs {err['mean_map_baseline']:.3f}")
    assert err["radial_density_mae"]["surrogate"] < 0.7 * err["radial_density_mae"]["mean_baseline"]
    assert np.isfinite(err["surrogate_top"]) and err["surrogate_validation"] < 1.05 * err["mean_map_baseline"]
    # only re-simulated sets whose verified error is within the bar count as resolved
    assert out["bar"]["contact_map_error"] == 0.05
    assert 0 <= out["resolved_verified"] <= out["verified"] == 32
    assert out["value"] == pytest.approx(out["resolved_verified"] * 3600.0 / out["seconds"]["total"])
    assert (out["resolved_verified"] > 0) == (min(out["top_errors"]) <= 0.05)
    assert out["scored_per_hr"] > out["value"]
