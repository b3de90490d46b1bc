"""CLI front end and the routing cost model on the CPU: the reference's
dry-run and explain lines, and the non-negative fit recovering known
coefficients."""
import numpy as np
import pytest

from film_grain_amd import __main__ as cli
from film_grain_amd import host, route


@pytest.fixture()
def image(tmp_path):
    rng = np.random.default_rng(0)
    rgb8 = (rng.random((24, 40, 3)) * 255).astype(np.uint8)
    path = tmp_path / "in.ppm"
    host.save_rgb8(str(path), rgb8, "ppm")
    return path


def test_dry_run_lines(image, tmp_path, capsys):
    rc = cli.main([str(image), "-o", str(tmp_path / "out.png"), "--dry-run", "--explain", "--device", "cpu",
                   "--zoom", "2", "--iters", "100"])
    out = capsys.readouterr().out.splitlines()
    assert rc == 0
    # src/choose.rs: n_samples > 96 routes auto to the grain-wise integrator
    assert out[0] == "dry-run: Grain algorithm, input 40x24 → output 80x48, samples 100"
    assert out[1].startswith("sigma/mean 0.000, rm/mean ") and out[1].endswith("samples 100 on Cpu")


def test_invalid_arguments_report_the_reference_error(image, tmp_path, capsys):
    rc = cli.main([str(image), "-o", str(tmp_path / "o.png"), "--dry-run", "--radius", "-1"])
    assert rc == 1
    assert capsys.readouterr().err.startswith("error: ")


def test_probe_crop_is_centred_and_bounded():
    assert route.crop_box(1920, 1080) == (448, 28, 1472, 1052)
    assert route.crop_box(300, 200) == (0, 0, 300, 200)
    x0, y0, x1, y1 = route.crop_box(513, 4096, side=512)
    assert (x1 - x0, y1 - y0) == (512, 512) and y0 == (4096 - 512) // 2
    rgb = np.zeros((40, 60, 3), np.float32)
    args = host.CliArgs(roi="4,3:50,33")
    assert route.roi_crop(args, rgb).shape == (30, 46, 3)
