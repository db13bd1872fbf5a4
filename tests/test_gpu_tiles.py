"""Thread-block tiles and XCD slab orders of the stencil kernels
(fcfd_set_tile) and the tile autotuner: every tile and order gives the same
step bit for bit (a cell's arithmetic does not depend on them), shapes
outside the supported set are refused, and the tuner leaves a valid choice
on every kernel without changing the run."""
import numpy as np
import pytest

from cases import bubble, case2d, case3d_odd

pytestmark = pytest.mark.gpu

FIELDS = ("phi", "u", "v", "p", "nu", "nv")


def _run(case, tile, steps=3):
    from fortrancfd_amd import Solver
    from fortrancfd_amd.types import TILE_KERNELS
    s = Solver(case)
    if tile is not None:
        for k in TILE_KERNELS:
            s.set_tile(k, *tile)
    s.push("phi", bubble(case, 0.0, 1))
    s.problem_init(False)
    for _ in range(steps):
        s.step()
    names = FIELDS + (("w", "nw") if case.dim == 3 else ())
    return {f: s.pull(f) for f in names}


@pytest.mark.parametrize("mk", [lambda: case2d(32, ppe_tol=1e-12), lambda: case3d_odd(ppe_tol=1e-12)],
                         ids=["2d", "3d_odd"])
def test_every_tile_gives_the_same_step(mk):
    from fortrancfd_amd.autotune import TILES
    case = mk()
    ref = _run(case, None)
    # the tuner's candidates and slabs that leave a partial last slab
    for t in list(TILES) + [(64, 4, 3), (128, 1, 5), (64, 2, 7)]:
        got = _run(case, t)
        for f in ref:
            assert np.array_equal(got[f], ref[f]), (t, f)


def test_tile_shapes_are_checked():
    from fortrancfd_amd import Solver
    from fortrancfd_amd.solver import FcfdError
    s = Solver(case2d(16))
    assert s.get_tile(0) == (64, 4, 0)
    for bad in ((32, 8, 0), (64, 8, 0), (512, 1, 0), (96, 2, 0), (128, 4, 0), (64, 4, -1), (64, 4, 5000)):
        with pytest.raises(FcfdError):
            s.set_tile(0, *bad)
    with pytest.raises(FcfdError):
        s.set_tile(5, 64, 4)
    s.set_tile(4, 256, 1, 3)
    assert s.get_tile(4) == (256, 1, 3)


def test_tile_tuner_keeps_the_run():
    from fortrancfd_amd import Solver
    from fortrancfd_amd.autotune import TILES, tune_tiles
    from fortrancfd_amd.types import TILE_KERNELS
    case = case3d_odd(ppe_tol=1e-12)
    s = Solver(case)
    s.push("phi", bubble(case, 0.0, 1))
    s.problem_init(False)
    res = tune_tiles(s)
    steps = len(TILES) + 1
    for k, region in TILE_KERNELS.items():
        assert tuple(res["best"][region]) in set(TILES)
        assert s.get_tile(k) == tuple(res["best"][region])
        assert len(res["us"][region]) >= 1
    # the tuning steps are ordinary steps: same state as untuned steps
    ref = _run(case, None, steps=steps)
    for f in ref:
        assert np.array_equal(s.pull(f), ref[f]), f
