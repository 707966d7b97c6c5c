"""Shared test set-up: the `gpu` marker, the package loader (the package lives in
the directory binary-supernova_amd/, imported as binary_supernova_amd) and the
oracle import path."""
from __future__ import annotations

import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def load_package():
    name = "binary_supernova_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = ROOT / "binary-supernova_amd"
    spec = importlib.util.spec_from_file_location(name, pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libbsn_hip.so")
