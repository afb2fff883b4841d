"""The grouped npz file (no GPU): groups, records with arrays, scalars,
strings, tuples and lists of records, replacement of one group, and loading
without pickle."""

import numpy as np
import pytest

from steam_amd import files as F


def _record(rng):
    return {'h': rng.standard_normal((4, 3, 2)).astype(np.float32), 'nx': 4,
            'dx': 62.5, 'turbulon_shape': 'mexican_hat', 'bounds': (1.0, 2.0, 0.0, 0.03),
            'seed': None, 'periodic_x': True,
            'class_increments': [{'h': rng.standard_normal(3), 'k': 2000.0},
                                 {'h': rng.standard_normal(5), 'k': 1000.0}]}


def test_round_trip_and_groups(tmp_path):
    rng = np.random.default_rng(0)
    path = tmp_path / "run.npz"
    root = _record(rng)
    F.write_group(path, '/', root, replace_file=True)
    nest = {'h': np.ones((2, 2, 2), np.float32), 'x0': 250.0}
    F.write_group(path, 'refinements/r0', nest)
    assert F.groups(path) == ['/', 'refinements/r0']
    back = F.read_group(path, '/')
    np.testing.assert_array_equal(back['h'], root['h'])
    assert back['nx'] == 4 and back['dx'] == 62.5 and back['turbulon_shape'] == 'mexican_hat'
    assert tuple(back['bounds']) == root['bounds'] and back['seed'] is None
    assert back['periodic_x'] is True
    assert [c['k'] for c in back['class_increments']] == [2000.0, 1000.0]
    np.testing.assert_array_equal(back['class_increments'][1]['h'],
                                  root['class_increments'][1]['h'])
    assert F.read_group(path, 'refinements/r0')['x0'] == 250.0
    # a plain numpy load needs no pickle
    with np.load(path, allow_pickle=False) as data:
        assert 'root|h' in data.files


def test_rewriting_a_group_keeps_the_others(tmp_path):
    rng = np.random.default_rng(1)
    path = tmp_path / "run.npz"
    F.write_group(path, '/', _record(rng))
    F.write_group(path, 'refinements/r0', {'a': np.arange(3)})
    shorter = dict(_record(rng), class_increments=[{'h': np.zeros(1), 'k': 1.0}])
    F.write_group(path, '/', shorter)
    back = F.read_group(path, '/')
    assert len(back['class_increments']) == 1
    with np.load(path) as data:
        assert not any(k.startswith('root/class_increments.1') for k in data.files)
    np.testing.assert_array_equal(F.read_group(path, 'refinements/r0')['a'], np.arange(3))
    with pytest.raises(KeyError):
        F.read_group(path, 'refinements/r9')
