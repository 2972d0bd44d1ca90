"""The frame-loop driver (pyro_amd/run.py) on the GPU: plain V-cycles, a
tolerance solver, and a run resumed from its own checkpoint."""
import pytest

from harness import MINI_SCENE
from pyro_amd import run

pytestmark = pytest.mark.gpu


def test_driver_runs_and_resumes(tmp_path, capsys):
    scene = tmp_path / "mini.toml"
    scene.write_text(MINI_SCENE)
    ck = tmp_path / "state.ckpt"
    assert run.main([str(scene), "--duration", "0.1", "--preroll", "0", "--save-state", str(ck),
                     "--slice", "--out", str(tmp_path)]) == 0
    out = capsys.readouterr().out
    assert "--- averages" in out and "substeps" in out and (tmp_path / "slice.png").exists()
    assert run.main([str(scene), "--duration", "0.2", "--preroll", "0", "--load-state", str(ck),
                     "--solver", "mgpcg", "--tol", "1e-5", "--quiet"]) == 0
    out = capsys.readouterr().out
    assert "mean_solver_iterations" in out and "'frames': 3" in out


def test_learned_weights_made_on_the_command_line_drive_the_loop(tmp_path, capsys):
    """`python -m pyro_amd.learned --out W` trains the pool network and saves
    it; learned.trained(W) then loads that file (no second training), and the
    driver runs with --solver learned --weights W."""
    import numpy as np

    from pyro_amd import learned

    w = tmp_path / "unet.npz"
    learned.main(["--out", str(w), "--frames", "24", "--steps", "40"])
    assert w.exists() and "saved to" in capsys.readouterr().out
    a, b = learned.load(w), learned.trained(w)
    assert np.array_equal(a.params(), b.params()) and a.output_scale == b.output_scale
    scene = tmp_path / "mini.toml"
    scene.write_text(MINI_SCENE)
    assert run.main([str(scene), "--duration", "0.1", "--preroll", "0", "--solver", "learned",
                     "--weights", str(w), "--quiet"]) == 0
    assert "mean_solver_iterations" in capsys.readouterr().out
