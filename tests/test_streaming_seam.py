"""The streamed driver's signature against the reference's call (no GPU):
``simulate`` calls ``streamed_cascade_loop`` with the plan and the scratch
directory positional and the rest by keyword
(steam/simulate.py:1380-1404; signature steam/streaming.py:850-871)."""

import inspect

import pytest


def _reference_call_args():
    positional = ["h_profile", "qt_profile", "z_profile", "grids", "C_h_k", "C_qt_k",
                  "b_h_k", "b_qt_k", "aspect_k", "h_min", "h_max", "qt_min", "qt_max",
                  "min_distance_to_ground", "sparsity_factors", "child_seeds", "plan",
                  "scratch_dir"]
    keywords = dict(n_scale_classes_per_dyad=1, turbulon_shape='mexican_hat', device='cuda',
                    comp_k=None, increment_store=None, completed=None, _crash_hook=None)
    return positional, keywords


def test_reference_call_binds():
    from steam_amd.streaming import streamed_cascade_loop
    positional, keywords = _reference_call_args()
    bound = inspect.signature(streamed_cascade_loop).bind(*positional, **keywords)
    assert bound.arguments['plan'] == 'plan'
    assert bound.arguments['scratch_dir'] == 'scratch_dir'
    assert bound.arguments['class_seeds'] == 'child_seeds'
    # the reference's other keywords bind too
    inspect.signature(streamed_cascade_loop).bind(
        *positional, flux_noise_scale=0.2, zero_bottom=True, zero_top=True)


@pytest.mark.parametrize("name", ["increment_store", "completed", "_crash_hook"])
def test_disk_features_are_refused_by_name(name):
    """The on-disk store's features raise NotImplementedError (not a
    TypeError from the signature), before anything touches the device."""
    from steam_amd.streaming import streamed_cascade_loop
    positional, keywords = _reference_call_args()
    keywords[name] = object()
    with pytest.raises(NotImplementedError, match=name):
        streamed_cascade_loop(*positional, **keywords)


def test_resident_store_is_a_no_op_handle():
    from steam_amd.streaming import ResidentStore
    st = ResidentStore("/tmp/unused")
    assert st.scratch_dir == "/tmp/unused"
    st.destroy()
    assert st.destroyed
