"""Loads the library's Python binding (interfacial-tension-operators-lbm_amd/__init__.py) by path."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "interfacial-tension-operators-lbm_amd")


def load():
    if "itolbm" in sys.modules:
        return sys.modules["itolbm"]
    spec = importlib.util.spec_from_file_location("itolbm", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["itolbm"] = mod
    spec.loader.exec_module(mod)
    return mod
