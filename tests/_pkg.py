"""Import the package directory (its name is not a Python identifier)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "machine-learning-based-implicit-physical-parametric-inference-for-cfd-simulation-promotion_amd")


def load():
    if "ippi_amd" in sys.modules:
        return sys.modules["ippi_amd"]
    spec = importlib.util.spec_from_file_location("ippi_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["ippi_amd"] = m
    spec.loader.exec_module(m)
    return m
