"""MI355X-native level-set DEM timestep (FunDEM / kaiqideng/LSDEM2.0 hot path).

The directory name is not a Python identifier; load it with `load_package()` from
`__graft_entry__` (importlib with an explicit location), which registers it as `lsdem_amd`.
"""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
