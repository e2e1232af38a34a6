"""Pass window geometry and vector grids (host only): the pass schedule of
WorkPIV against hand-computed window sizes and the reference's error texts,
and the grid rule against hand counts."""

import numpy as np
import pytest

from fluidimage_amd.works.piv import schedule


def test_pass_schedule_true_division_then_int():
    # 64 / 3 = 21.33 -> 21 px, then 21.33 / 2 = 10.67 -> 10 px
    sizes = schedule.pass_schedule(64, None, 3, [3, 2])
    assert [schedule.window_shapes(*s)[0] for s in sizes] == [(64, 64), (21, 21), (10, 10)]
    sizes = schedule.pass_schedule((64, 48), (32, 32), 2, 2)
    assert [schedule.window_shapes(*s) for s in sizes] == [((64, 48), (32, 32)),
                                                            ((32, 24), (16, 16))]


def test_schedule_errors_are_the_references():
    with pytest.raises(ValueError, match="coeff_zoom has to be an integer"):
        schedule.pass_schedule(64, None, 3, [2])
    with pytest.raises(NotImplementedError, match="shape_crop_im0 has to be one or two"):
        schedule.window_shapes((1, 2, 3))
    with pytest.raises(NotImplementedError, match="shape_crop_im1 has to be one or two"):
        schedule.window_shapes(32, "x")
    with pytest.raises(ValueError, match="inferior or equal"):
        schedule.window_shapes(32, (40, 16))


def test_schedule_params_for_uneven_schedules():
    from fluidimage_amd.params import Params

    p = Params()
    p._set_child("piv0", {"shape_crop_im0": 48, "shape_crop_im1": None})
    p._set_child("multipass", {"number": 1, "coeff_zoom": 2})
    for windows in ((128, 64, 32), (64, 32, 16), (48, 24)):
        schedule.schedule_params(p, windows)
        assert p.multipass.number == len(windows)
    with pytest.raises(ValueError, match="not a chain"):
        schedule.schedule_params(p, (64, 40))


@pytest.mark.parametrize("length, window, overlap, expected", [
    # 2048 px, 32 px windows at 50 %: centres 16, 32, ..., 2032 (127)
    (2048, 32, 0.5, 127),
    # 2048 px, 16 px: 8 .. 2032 (254) then 2040 (gap 8 > 8 // 2.5 = 3)
    (2048, 16, 0.5, 255),
    # 512 px, 48 px windows, step 24: 24 .. 480 (20); the gap 8 is not > 24 // 2.5 = 9
    (512, 48, 0.5, 20),
])
def test_grid_positions_counts(length, window, overlap, expected):
    step = schedule.grid_step(window, overlap)
    pos = schedule.grid_positions(length, window // 2, window // 2 + window % 2, step)
    assert len(pos) == expected
    assert pos[0] == window // 2 and pos[-1] <= length - window // 2


def test_workpiv_passes_and_grid_on_the_host():
    """WorkPIV builds one work per pass with the schedule's windows; its
    grids are row-major (ivec = iy_index nx + ix_index). Plans without
    displacement_max need no GPU."""
    from fluidimage_amd.works.piv.multipass import WorkPIV

    p = WorkPIV.create_default_params()
    p.piv0.shape_crop_im0 = 64
    p.multipass.number = 3
    w = WorkPIV(p)
    assert [wk.shape_crop_im0 for wk in w.works_piv] == [(64, 64), (32, 32), (16, 16)]
    assert [wk.index_pass for wk in w.works_piv] == [0, 1, 2]
    w._prepare_with_image(imshape=(256, 320))
    last = w.works_piv[-1]
    assert last.nb_vectors == len(last.iyvecs) * len(last.ixvecs)
    assert np.array_equal(last.ixvecs_grid[: len(last.ixvecs)], last.ixvecs)
    assert np.all(last.iyvecs_grid[: len(last.ixvecs)] == last.iyvecs[0])
