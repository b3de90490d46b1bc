"""The tuned launch table's plumbing on the host: bucket keys from the
library (fg_launch_bucket), the JSON table's format, and the built-in table
header generated from it.  Which shapes the table picks, and that they change
only speed, is checked on the GPU (test_gpu_tuning.py)."""
import json

import pytest

from film_grain_amd import autotune, build, capi


def _desc(w=1920, h=1080, zoom=1.0, n=64, dist=0, rm=0.2, delta=0.1):
    return capi.make_desc(w, h, int(w * zoom), int(h * zoom), n, dist, 7, zoom, delta, rm, 0.1, -2.4, 0.5)


def test_buckets_follow_the_documented_quantisation():
    key = capi.launch_bucket(_desc(), "pixel")
    assert (key.algo, key.dist_kind, key.zoom_q, key.n_bucket) == (0, 0, 0, 64)
    assert key.ratio_h == 2  # rm / delta = 2: one octave
    # plane size does not enter; N rounds up to a power of two
    assert capi.bucket_name(capi.launch_bucket(_desc(w=128, h=333), "pixel")) == capi.bucket_name(key)
    assert capi.launch_bucket(_desc(n=33), "pixel").n_bucket == 64
    assert capi.launch_bucket(_desc(n=65), "pixel").n_bucket == 128
    # zoom in quarter octaves, the ratio in half octaves
    assert capi.launch_bucket(_desc(zoom=2.0), "pixel").zoom_q == 4
    assert capi.launch_bucket(_desc(zoom=1.19), "pixel").zoom_q == 1
    assert capi.launch_bucket(_desc(rm=0.4), "pixel").ratio_h == 4
    assert capi.launch_bucket(_desc(rm=0.141), "pixel").ratio_h == 1
    # grain-wise keys use rm * zoom; the law is part of the key
    g = capi.launch_bucket(_desc(zoom=2.0, rm=0.25), "grain")
    assert (g.algo, g.ratio_h) == (1, -2)
    assert capi.launch_bucket(_desc(dist=1), "pixel").dist_kind == 1
    assert capi.launch_bucket(_desc(rm=0.0), "pixel").ratio_h == -1000


def test_json_table_parses_and_names_known_fields():
    table = json.loads(build.TUNING_TABLE.read_text())
    assert table, "the tuned table is empty"
    for key, entry in table.items():
        algo, dist, zq, rh, nb = build.parse_bucket(key)
        assert nb & (nb - 1) == 0 and algo in (0, 1) and dist in (0, 1)
        assert set(entry["config"]) <= set(build.LAUNCH_FIELDS), key
    entries = autotune.table_entries(table)
    assert [capi.bucket_name(k) for k, _ in entries] == sorted(table)


def test_builtin_header_holds_every_entry(tmp_path, monkeypatch):
    monkeypatch.setattr(build, "BUILD", tmp_path)
    text = build.launch_table_header().read_text()
    table = json.loads(build.TUNING_TABLE.read_text())
    assert f"kBuiltinLaunchCount = {len(table)};" in text
    for key in table:
        assert f"// {key}" in text


def test_bad_keys_are_refused():
    for bad in ("pixel|const|z1|rh0|N64", "tile|const|zq0|rh0|N64", "pixel|const|zq0|rh0"):
        with pytest.raises(ValueError):
            build.parse_bucket(bad)
