"""Autotuner cache on the host (no GPU): load, lookup by explicit device arch, save/reload
round trip, and rejection of malformed files."""
import pytest

from fft_demo_amd import tune


def _entry_line(arch, nx, ny, batch, desc, split, limit, ms, dms, cand):
    return (f"{arch} {nx} {ny} {batch} {desc.precision} {desc.in_fmt} {desc.out_fmt} {desc.mid_fmt} "
            f"{desc.direction} {desc.in_order} : {split} {limit} {ms:.6f} {dms:.6f} {cand}\n")


def test_cache_load_lookup_save_roundtrip(tmp_path):
    from fft_demo_amd.plan import make_desc

    tune.clear()
    d = make_desc(precision="split_fp16")
    f = tmp_path / "tune.txt"
    f.write_text("mfft-tune 1\n" + _entry_line("gfx950", 65536, 0, 256, d, 64, 67108864, 0.081, 0.094, 30) +
                 _entry_line("gfx950", 1024, 1024, 64, d, 0, 0, 0.5, 0.5, 5))
    tune.load(f)
    e = tune.lookup((65536,), 256, arch="gfx950", precision="split_fp16")
    assert e == {"split_hint": 64, "workspace_limit": 67108864, "ms": pytest.approx(0.081),
                 "default_ms": pytest.approx(0.094), "candidates": 30}
    assert tune.lookup((65536,), 256, arch="gfx942", precision="split_fp16") is None
    assert tune.lookup((65536,), 128, arch="gfx950", precision="split_fp16") is None  # batch is in the key
    assert tune.lookup((65536,), 256, arch="gfx950", precision="fp32") is None  # precision is in the key
    assert tune.lookup((1024, 1024), 64, arch="gfx950", precision="split_fp16")["ms"] == pytest.approx(0.5)
    g = tmp_path / "again.txt"
    tune.save(g)
    tune.clear()
    assert tune.lookup((65536,), 256, arch="gfx950", precision="split_fp16") is None
    tune.load(g)
    assert tune.lookup((65536,), 256, arch="gfx950", precision="split_fp16")["split_hint"] == 64
    tune.clear()


def test_cache_rejects_malformed(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("not a cache\n")
    with pytest.raises(ValueError):
        tune.load(bad)
    bad.write_text("mfft-tune 1\ngfx950 1 2 3\n")
    with pytest.raises(ValueError, match="malformed"):
        tune.load(bad)
    with pytest.raises(ValueError):
        tune.load(tmp_path / "missing.txt")
