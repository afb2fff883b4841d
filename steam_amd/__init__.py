"""steam_amd — the STEAM field-generation hot path on AMD MI355X (gfx950).

World-keyed Philox4x32-10 noise with CMS extremal-Levy variates, turbulon
kernel synthesis, the periodic-x/y zero-padded-z FFT overlap-add convolution,
the multiplicative flux-cascade advance and the scalar increments of the
cascade, as hand-written HIP kernels (libsteam_hip.so) behind the reference's
Python API:

    steam_amd.noise       keyed_uniforms, class_key, center_indices
    steam_amd.turbulons   SHAPES, mexican_hat, morlet_omega0_6, power_law
    steam_amd.utils       convolve_fft_xy_oa_z, fold_kernel_to_field,
                          zoom_trilinear
    steam_amd.simulate    CONVOLVE, NoiseRegion, _keyed_sparse_levy,
                          _levy_from_uniforms, _turbulon_envelope,
                          _compute_all_grids, _advance_flux,
                          simulate_flux_only
    steam_amd.scalars     advective_weight, joint_pattern, product_norm,
                          bounded_amplitude_add
    steam_amd.cascade     cascade_loop, project_onto_bounds, compose_output,
                          compose_flux_output, simulate_fields
    steam_amd.api         simulate, refine (the reference's signatures; the run
                          written to a grouped npz file), read_ledger; also
                          reached as steam_amd.simulate.simulate / .refine
    steam_amd.files       the grouped npz file (NetCDF-style groups)
    steam_amd.params      _compute_normalization, _bound_buffers,
                          _compensation_profiles, root_setup
    steam_amd.ledger      MeanReduction, FluxAdvanceLedger, PatternLedger,
                          BoundedAddLedger, ProjectionLedger, RunLedger
    steam_amd.streaming   tile geometry, tiled regrid, tiled flux cascade
    steam_amd.batch       many uneven realisations per launch
    steam_amd.serve       continuous-batching tile server
"""

__version__ = "0.1.0"

from . import _lib  # noqa: F401  (binds torch's HIP runtime first)


def library_version():
    return _lib.load().steam_version()


def cascade_loop(*args, **kwargs):
    """steam.cascade_loop (steam/__init__.py:6): see steam_amd.cascade."""
    from .cascade import cascade_loop as _cascade_loop
    return _cascade_loop(*args, **kwargs)
