"""The C-ABI library loads, exports every declared symbol, and rejects bad
parameters with ccp's messages before touching a GPU."""
import os
import re

import numpy as np
import pytest

import tfm_amd
from tfm_amd import ccp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(ccp_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    from tfm_amd import surrogate

    L = ccp.lib()
    syms = declared_symbols("ccpmi.h")
    assert len(syms) >= 30
    assert syms == set(ccp.EXPORTS)
    nn = declared_symbols("ccpmi_nn.h")
    assert nn == set(surrogate.NN_EXPORTS)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["ccpmi.h", "ccpmi_nn.h"]
    for s in syms | nn:
        assert hasattr(L, s), s


def test_oracle_not_linked_into_product():
    # the product library must not depend on the oracle
    data = open(ccp.LIB_PATH, "rb").read()
    assert b"orc_" not in data and b"ccp_oracle" not in data


BASE = {"number_of_particles": 1024, "nucleus_radius": 10.53, "seed": 1}


@pytest.mark.parametrize("change,msg", [
    ({"nucleus_radius": 0}, "nucleus_radius out of range"),
    ({"nucleus_radius": -3}, "nucleus_radius out of range"),
    ({"opening_radius": 11}, "opening_radius out of range"),
    ({"opening_radius": 2, "bleb_radius": 1}, "bleb_radius out of range"),
    ({"number_of_particles": 0}, "number_of_particles out of range"),
    ({"temperature": 1000}, "temperature out of range"),
    ({"number_of_lbs": 100000}, "number_of_lbs out of range"),
    ({"frames_per_file": 0}, "frames_per_file out of range"),
    ({"nucleus_radius": 7.0}, "chromatin volume fraction above 0.4"),
    ({"number_of_lbs": 900}, "lbs area fraction above 0.4"),
    ({"steps_per_frame": 10000}, "steps_per_frame out of range"),
    ({"threads_per_block": 1025}, "threads_per_block out of range"),
    ({"force_field": "xyz"}, "force_field must be ccp or lj"),
    ({"cell_size": 1.5}, "cell_size below the interaction cutoff"),
    ({"chain_lengths": "500,500"}, "chain_lengths must sum to N"),
])
def test_parameter_errors(change, msg):
    p = dict(BASE)
    p.update(change)
    with pytest.raises(tfm_amd.CcpError) as e:
        tfm_amd.Simulation(p)
    assert str(e.value) == msg


def test_missing_radius_default_is_error():
    with pytest.raises(tfm_amd.CcpError, match="nucleus_radius out of range"):
        tfm_amd.Simulation({"number_of_particles": 100})


def test_first_occurrence_of_a_key_wins():
    # parmap inserts with std::map::insert, keeping the first value (util.cu:72-80)
    with pytest.raises(tfm_amd.CcpError, match="nucleus_radius out of range"):
        tfm_amd.Simulation("number_of_particles 1024\nnucleus_radius 0\nnucleus_radius 10.5\n")


def test_sequence_pattern_is_host_only_and_deterministic():
    from tfm_amd.ccp import sequence_pattern

    p = {"number_of_particles": 4096, "nucleus_radius": "15.0", "force_field": "lj", "het_fraction": 0.3,
         "lad_fraction": 0.15, "seed": 11, "number_of_chains": 2}
    a = sequence_pattern(p, 4096)
    assert np.array_equal(a, sequence_pattern(p, 4096))
    assert abs(((a & 1) > 0).mean() - 0.3) < 0.06
    assert ((a & 2) > 0).sum() == 2 * round(0.15 * 2048)
    assert not np.array_equal(a, sequence_pattern({**p, "seed": 12}, 4096))
    # replicas keep the pattern
    assert np.array_equal(a, sequence_pattern({**p, "replica": 3}, 4096))
    with pytest.raises(tfm_amd.CcpError):
        sequence_pattern(p, 100)


def test_ensemble_draws():
    # chain lengths log-uniform in range, parameters in their intervals, seeds
    # consecutive, the same draw for the same seed
    from tfm_amd.ensemble import draw_systems

    a = draw_systems(64, 512, 16384, 4000)
    assert a == draw_systems(64, 512, 16384, 4000)
    n = np.array([s["n_beads"] for s in a])
    assert n.min() >= 512 and n.max() <= 16384 and np.median(n) < (512 + 16384) / 2
    for s, i in zip(a, range(4000, 4064)):
        assert s["seed"] == i
        assert 0.1 <= s["volume_fraction"] <= 0.35 and 0.2 <= s["eps_het"] <= 1.0 and 0.2 <= s["eps_lam"] <= 1.5


def test_ensemble_layout_fits_the_kernels_index_arithmetic():
    """The 256-system ensemble (512-16384 beads, seeds 4000 + i) laid out on
    the host at both kernel shapes: every system's cells are a contiguous
    range of the shared cell range sized by its own grid (cps^3 with cps =
    2 ceil(R_n / side), at least 4), the cell indices' byte offsets (index x 4)
    and the candidate list's slot offsets (slot x 16) fit the 32-bit offsets
    of the force kernels, the order pass's 256-cell blocks cover the range,
    and the bucket stays addressable at the largest capacity growth allows
    (4096 entries per cell)."""
    import math

    from tfm_amd.ensemble import draw_systems

    systems = draw_systems(256, 512, 16384, 4000)
    n_total = sum(s["n_beads"] for s in systems)
    for shape in ({"lanes_per_bead": 3, "cell_reach": 1, "cell_size": 2.5},
                  {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.375}):
        lay = tfm_amd.batch_layout(systems, {"timestep": 0.001, **shape})
        assert lay["beads"] == n_total
        off = lay["cell_offsets"].astype(np.int64)
        assert off[0] == 0 and off[-1] == lay["cells"]
        side = shape["cell_size"]
        for s, d in enumerate(systems):
            rco = float(np.float32(1.154701))
            R_n = float(np.float32(float(f"{rco / 2 + rco / 2 * (d['n_beads'] / d['volume_fraction']) ** (1 / 3):.6f}")))
            cps = max(2 * math.ceil(np.float32(R_n) / np.float32(side)), 4)
            assert off[s + 1] - off[s] == cps ** 3, s
        assert lay["cells"] * 4 < 2 ** 32
        assert n_total * 16 < 2 ** 32
        blocks = (lay["cells"] + 255) // 256
        assert blocks * 256 >= lay["cells"] and blocks < 2 ** 31
        assert lay["bucket_bytes"] == lay["cells"] * 32 * 16
        assert lay["cells"] * 4096 * 16 < 2 ** 63


def test_half_cell_batches_above_the_tested_size_are_refused():
    # eight lanes over half-cutoff cells faulted the GPU once on a 1.1M-bead
    # ensemble with no cause found: the library refuses that shape above
    # 524288 beads and the batch autotuner leaves it out
    from tfm_amd import autotune
    from tfm_amd.ensemble import draw_systems

    systems = draw_systems(256, 512, 16384, 4000)
    half = {"timestep": 0.001, "lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25}
    with pytest.raises(tfm_amd.CcpError, match="limited to 524288 beads per batch"):
        tfm_amd.batch_layout(systems, half)
    small = [dict(s, n_beads=2048) for s in systems]
    assert tfm_amd.batch_layout(small, half)["beads"] == 524288
    shapes = autotune.candidates(2.5, sum(s["n_beads"] for s in systems))
    assert {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25} not in shapes and len(shapes) == 5
    assert len(autotune.candidates(2.5, 524288)) == 6
