"""Shared pytest set-up: the `gpu` marker and loaders for the product package
(whose directory name is not an importable identifier) and the oracle."""
import importlib.util
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "c-version-lithography_amd"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def _load(name: str, path: Path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def litho_amd():
    return _load("litho_amd", PKG / "litho_amd.py")


def targets():
    return _load("litho_targets", PKG / "targets.py")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    la = litho_amd()
    ctx = la.Context(0)
    yield ctx
    ctx.close()
