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
