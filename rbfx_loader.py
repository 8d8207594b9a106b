"""Import helper: the package directory is named radialbasisfunctions.jl_amd, which is
not a valid Python identifier, so it is loaded from its path and registered as
`radialbasisfunctions_jl_amd`."""
import importlib.util
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "radialbasisfunctions.jl_amd"
NAME = "radialbasisfunctions_jl_amd"


def load():
    if NAME in sys.modules:
        return sys.modules[NAME]
    spec = importlib.util.spec_from_file_location(NAME, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[NAME] = mod
    spec.loader.exec_module(mod)
    return mod
