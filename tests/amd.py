"""Loads the product package (its directory name is not a Python identifier)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "labinius-fast-challenge_amd")


def load():
    if "labinius_amd" in sys.modules:
        return sys.modules["labinius_amd"]
    spec = importlib.util.spec_from_file_location("labinius_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["labinius_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
