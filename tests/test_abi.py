"""The ctypes structures of the Python binding have the C layout the header
declares (sizes and field offsets, checked by compiling the header with the
host C compiler).  No GPU needed."""

import ctypes
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

STRUCTS = {
    "steam_region": ("Region", ["key", "origin", "world_shape", "sparsity"]),
    "steam_noise_params": ("NoiseParams", ["region", "alpha", "per_class_scale",
                                           "shift", "n_zero", "zero_bottom",
                                           "zero_top"]),
    "steam_conv_info": ("ConvInfo", ["n_fft", "n_blocks", "block_len",
                                     "workspace_bytes", "spectrum_bytes",
                                     "spectral_elems"]),
    "steam_flux_ledger": ("FluxLedger", ["entering_total", "entering_count",
                                         "noise_abs_total", "noise_abs_count",
                                         "realized_total", "realized_count",
                                         "n_clipped", "rescale", "flatten"]),
    "steam_zdiff": ("ZDiff", ["uniform", "two_dz", "dz_first", "dz_last", "a_dev",
                              "b_dev", "c_dev"]),
    "steam_bounded_add_ledger": ("BoundedAddLedgerC", ["a0", "demean", "scale", "n_loop",
                                                       "n_scale", "final_demean", "mu"]),
    "steam_batch_item": ("BatchItem", ["plan", "spectrum_dev", "workspace_dev",
                                       "flux_dev", "noise_dev", "sk_dev", "noise",
                                       "window", "sigma_in", "flags", "ledger"]),
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             f'#include "{os.path.join(ROOT, "include", "steam_hip.h")}"',
             'int main(void) {']
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src = d / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "abi"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        name, field, value = line.split()
        layout[(name, field)] = int(value)
    return layout


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_ctypes_layout_matches_header(cname, c_layout):
    from steam_amd import _lib
    pyname, fields = STRUCTS[cname]
    cls = getattr(_lib, pyname)
    assert ctypes.sizeof(cls) == c_layout[(cname, "size")]
    for f in fields:
        assert getattr(cls, f).offset == c_layout[(cname, f)], f
