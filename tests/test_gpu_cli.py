"""The command line on the device: a file render (the whole frame on the
device) equals the library's per-plane route, and --explain reports both
integrators' launch shapes and predicted times."""
import numpy as np
import pytest

from film_grain_amd import __main__ as cli
from film_grain_amd import host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("extra", [["--color-mode", "rgb", "--zoom", "2", "--iters", "64", "--algo", "pixel"],
                                   ["--color-mode", "luma", "--iters", "100"],
                                   ["--radius-dist", "lognorm", "--radius-stddev", "0.04", "--roi", "4,3:50,33"]])
def test_cli_render_equals_per_plane_route(tmp_path, capsys, extra):
    rng = np.random.default_rng(len(extra))
    rgb8 = (rng.random((36, 60, 3)) * 255).astype(np.uint8)
    src = tmp_path / "in.ppm"
    host.save_rgb8(str(src), rgb8, "ppm")
    out = tmp_path / "out.ppm"
    assert cli.main([str(src), "-o", str(out), "--explain", *extra]) == 0
    lines = capsys.readouterr().out.splitlines()
    print("\n".join(lines), flush=True)
    assert lines[0].startswith("rendered ")
    assert lines[2].startswith("hip: pixel-wise [") and "grain-wise [" in lines[2] and "ms/frame" in lines[2]
    got = host.load_image(str(out))
    a = cli._parser().parse_args([str(src), "-o", str(out), *extra])
    args = host.CliArgs(radius_dist=a.radius_dist, radius=a.radius, radius_stddev=a.radius_stddev, zoom=a.zoom,
                        iters=a.iters, algo=a.algo, color_mode=a.color_mode, roi=a.roi, seed=a.seed)
    want8, _, _ = host.render_rgb(args, host.load_image(str(src)), with_planes=False)
    assert np.array_equal((got * 255.0 + 0.5).astype(np.uint8), want8)
