"""Registers the package directory (its name is not a Python identifier) as u55c_statevector_30q_amd."""
import importlib.util
import os
import sys

NAME = "u55c_statevector_30q_amd"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load():
    if NAME in sys.modules:
        return sys.modules[NAME]
    d = os.path.join(ROOT, "u55c-statevector-30q_amd")
    spec = importlib.util.spec_from_file_location(NAME, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[NAME] = mod
    spec.loader.exec_module(mod)
    return mod
