"""MI355X-native TMS E-field FEM (the cuNIBS hot path) over a C-ABI HIP library.

Public names mirror the reference package (src/cunibs/__init__.py:5-35). They are resolved on
first access, so importing a submodule such as ``cunibs_amd.synthetic`` does not load the HIP
library; anything that needs the library raises if it is missing (there is no CPU fallback).
"""

from __future__ import annotations

_EXPORTS = {
    "Subject": "cunibs_amd.simulation",
    "Placement": "cunibs_amd.simulation",
    "FieldResult": "cunibs_amd.simulation",
    "Coil": "cunibs_amd.simulation",
    "RECOVERY_MODES": "cunibs_amd.simulation",
    "HeadMesh": "cunibs_amd.mesh",
    "load_mesh": "cunibs_amd.mesh",
    "ConductivityUQConfig": "cunibs_amd.uq",
    "ConductivityUQResult": "cunibs_amd.uq",
}

__all__ = sorted(_EXPORTS)


def __getattr__(name: str):
    if name in _EXPORTS:
        import importlib
        return getattr(importlib.import_module(_EXPORTS[name]), name)
    raise AttributeError(f"module 'cunibs_amd' has no attribute {name!r}")
