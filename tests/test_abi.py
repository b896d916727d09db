"""The C-ABI library loads and exports every symbol include/luzgpu.h declares (CPU)."""

import os
import re

from conftest import ROOT


def test_library_exports_header_symbols():
    from luz_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "luzgpu.h")).read()
    declared = set(re.findall(r"\b(luzgpu_[a-z_]+)\s*\(", header))
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        return
    from luz_amd import Context, _lib
    import pytest
    with pytest.raises(_lib.LuzGpuError):
        Context(0)
