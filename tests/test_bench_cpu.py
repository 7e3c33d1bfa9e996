"""bench.py reads its traffic and vector-ALU figures from the committed
rocprofv3 summaries under profiles/: they parse, and the figures are
plausible for the benchmark's force kernel (N = 131072, 100 algorithmic bytes
per bead in stage 1, between 8 and 16 beads per wavefront)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_profile_summaries_parse():
    b = load_bench()
    assert b.BYTES_FORCE_PER_BEAD == {1: 100, 2: 84}
    traffic = b.read_traffic(os.path.join(ROOT, "profiles", "k_force_typed_lj_n131072_traffic.txt"))
    algorithmic = b.BYTES_FORCE_PER_BEAD[1] * b.N
    # measured HBM bytes of a launch: at least the algorithmic bytes, and the
    # write-back of scattered bucket entries keeps it within a few times that
    assert traffic is not None and algorithmic <= traffic <= 4 * algorithmic
    valu = b.read_valu(os.path.join(ROOT, "profiles", "k_force_typed_lj_n131072_sq.txt"))
    assert valu is not None
    insts, us = valu
    # eight beads per wavefront (het beads) to sixteen (the others)
    assert 300 * (b.N // 16) <= insts <= 3000 * (b.N // 8)
    # the issue rate those imply stays below the 614.4 G/s peak
    assert 0.0 < insts / (us * 1e-6) / 1e9 < b.VALU_PEAK_GWIPS


def test_missing_summaries_read_as_none(tmp_path):
    b = load_bench()
    assert b.read_traffic(str(tmp_path / "absent.txt")) is None
    assert b.read_valu(str(tmp_path / "absent.txt")) is None
    p = tmp_path / "partial.txt"
    p.write_text("valu_insts_per_launch 5\n")
    assert b.read_valu(str(p)) is None
