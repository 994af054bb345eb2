"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import glob
import os
import re
import subprocess

import pytest

from conftest import REPO


def _declared():
    names = []
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(fadesim_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_header_symbols():
    from fadesim_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libfadesim.so not built (python -m fadesim_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b[TW] (\w+)", out))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing
    lib = _native.lib()
    assert lib.fadesim_version() >= 100
    for n in _declared():
        assert hasattr(lib, n)


def test_default_omega_rule():
    from fadesim_amd import _native
    w = _native.lib().fadesim_default_omega(128, 128)
    assert 1.97 < w < 1.99
    assert _native.lib().fadesim_default_omega(64, 64) < w


def test_shape_validation_without_gpu():
    import torch
    from fadesim_amd.solver import solve_batch
    with pytest.raises(ValueError):
        solve_batch(torch.ones((2, 48, 48)), 1.0, torch.zeros((2, 48)))
    with pytest.raises(ValueError):
        solve_batch(torch.ones((2, 64, 64)), 1.0, torch.zeros((2, 64)))  # not a GPU tensor
