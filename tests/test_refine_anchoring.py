"""Nest world anchoring and seed continuation (steam/simulate.py:4176-4270):
host arithmetic checked against hand-computed cells and the SeedSequence
spawn property.  No GPU needed."""

import numpy as np

from oracle import model as M


def _ladder():
    return 2000.0 / 2.0 ** np.arange(4)


def test_full_span_nest_anchors_at_the_world_origin():
    from steam_amd.simulate import _compute_all_grids, nest_world_anchoring
    k = _ladder()
    sph, zp = np.full(28, 100.), np.arange(28) * 30.
    world = _compute_all_grids(k, 2000., 2000., 800., (1, 1, 1), sph, zp)
    origins, shapes = nest_world_anchoring(k, 2000., 2000., 800., (1, 1, 1), sph, zp,
                                           0.0, 0.0, [0.0] * 4, [0.0] * 4,
                                           world['z_arrays'])
    assert origins == [(0, 0, 0)] * 4
    assert shapes == [(int(world['nx'][i]), int(world['ny'][i]), int(world['nz'][i]))
                      for i in range(4)]


def test_offset_nest_origin_is_the_nearest_world_cell():
    from steam_amd.simulate import _compute_all_grids, nest_world_anchoring
    k = _ladder()
    sph, zp = np.full(28, 100.), np.arange(28) * 30.
    world = _compute_all_grids(k, 2000., 2000., 800., (1, 1, 1), sph, zp)
    dx = np.asarray(world['dx'])
    pads = [2 * d for d in dx]            # two world cells of halo per class
    xmin, ymin = 600.0, 250.0
    z_nest = [np.asarray(z)[3:] for z in world['z_arrays']]   # starts at level 3
    origins, _ = nest_world_anchoring(k, 2000., 2000., 800., (1, 1, 1), sph, zp,
                                      xmin, ymin, pads, pads, z_nest)
    for i, (ox, oy, oz) in enumerate(origins):
        assert ox == int(round(xmin / dx[i])) - 2
        assert oy == int(round(ymin / float(world['dy'][i]))) - 2
        assert oz == 3


def test_seed_continuation_reproduces_the_root_stream():
    from steam_amd.simulate import nest_class_seeds
    root = np.random.SeedSequence(42).spawn(7)
    cont = nest_class_seeds(42, 4, 3)
    for a, b in zip(cont, root[4:]):
        assert M.class_key(a) == M.class_key(b)
    indep = nest_class_seeds(42, 4, 3, seed=5)
    assert M.class_key(indep[0]) == M.class_key(np.random.SeedSequence(5).spawn(3)[0])
