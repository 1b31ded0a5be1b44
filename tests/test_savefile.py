"""Save files in the reference's layout: text and binary round trips on the CPU, and (GPU) a run
continued from a saved frame against the uninterrupted run."""
import numpy as np
import pytest

from conftest import rel_err
from tissuefolding_paperversion_amd import configs, savefile


def _state(t, seed=0):
    rng = np.random.default_rng(seed)
    x = t.mesh.pos + 0.01 * rng.standard_normal(t.mesh.pos.shape)
    fg = np.tile(np.eye(3), (t.n_elements, 1, 1)) + 0.01 * rng.standard_normal((t.n_elements, 3, 3))
    gr = np.exp(t.rates * t.dt)
    return x, fg, gr


def test_frame_round_trip(tmp_path):
    t = configs.disc_tissue(radius=10.0, h=4.0)
    frames = [(_state(t, k), 60.0 * (k + 1)) for k in range(3)]
    with savefile.FrameWriter(str(tmp_path), t.mesh) as w:
        for (x, fg, gr), time in frames:
            w.write(time, x, fg, gr)
    got = savefile.read_frames(str(tmp_path))
    assert len(got) == 3
    for fr, ((x, fg, gr), time) in zip(got, frames):
        assert fr.time == time
        assert rel_err(fr.positions, x) < 1e-9                # 10 significant digits in the text
        assert np.array_equal(fr.fg, fg)                       # binary: exact
        assert np.array_equal(fr.growth_rate, gr)
        assert np.array_equal(fr.conn, t.mesh.conn)
        assert np.array_equal(fr.placement, t.mesh.node_placement)
        assert rel_err(fr.ref_positions, t.mesh.xref) < 1e-3  # 4 significant digits, as the reference writes


def test_frame_text_layout(tmp_path):
    """Column widths as the reference's streams write them (Simulation.cpp:9282-9355)."""
    t = configs.disc_tissue(radius=8.0, h=4.0)
    x, fg, gr = _state(t)
    with savefile.FrameWriter(str(tmp_path), t.mesh) as w:
        w.write(3600.0, x, fg, gr)
    lines = open(tmp_path / "Save_Frame").read().splitlines()
    assert lines[0] == "=============== TIME:       3600" + "=" * 51
    assert int(lines[1]) == t.n_nodes
    assert len(lines[2]) == 3 * 20 + 3 * 4
    el = lines[2 + t.n_nodes + 1]
    assert len(el) == 4 + 4 + 6 * 9 + 18 * 15
    assert lines[-1] == "=============== END OF TIME:       3600" + "=" * 44


def test_inconsistent_save_rejected(tmp_path):
    t = configs.disc_tissue(radius=10.0, h=4.0)
    other = configs.disc_tissue(radius=11.0, h=4.0)
    x, fg, gr = _state(t)
    with savefile.FrameWriter(str(tmp_path), t.mesh) as w:
        w.write(60.0, x, fg, gr)
    fr = savefile.read_frames(str(tmp_path))[0]

    class Dummy:
        pass
    with pytest.raises(ValueError):
        savefile.continue_from_save(Dummy(), fr, other.mesh)


@pytest.mark.gpu
def test_continue_from_save_matches_uninterrupted_run(tmp_path):
    from tissuefolding_paperversion_amd.solver import NewtonRaphsonSolver
    t = configs.wing_disc_tissue(a=25.0, b=16.0, h=4.0, z_levels=(0.0, 2.0, 6.0, 10.0), seed=3, growth_seed=4)
    t.rates *= 15.0
    a = NewtonRaphsonSolver(t)
    with savefile.FrameWriter(str(tmp_path), t.mesh) as w:
        for k in range(3):
            a.runOneStep(t.dt)
            fg, _ = a.get_fg()
            w.write(t.dt * (k + 1), a.positions, fg, np.exp(t.rates * t.dt))
    for _ in range(2):
        a.runOneStep(t.dt)
    fr = savefile.read_frames(str(tmp_path))[2]
    b = NewtonRaphsonSolver(t)
    savefile.continue_from_save(b, fr, t.mesh)
    for _ in range(2):
        b.runOneStep(t.dt)
    assert rel_err(b.positions, a.positions) < 1e-8
