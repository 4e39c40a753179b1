"""File names of the reference's 25 bundled coil models (src/cunibs/coil.py:10-35).

The files are the coil makers' dipole data and ship inside the reference package, not this one.
Each constant is the path of that file under ``COILS_ROOT``: the directory named by the
``CUNIBS_COILS_DIR`` environment variable, else the ``coils`` folder of an installed ``cunibs``
package when one is importable, else ``./coils``. ``Coil.load(coils.MAGSTIM_D70)`` then reads it
with :mod:`cunibs_amd.h5lite`.
"""
from __future__ import annotations

import importlib.util
import os
from pathlib import Path


def _root() -> Path:
    env = os.environ.get("CUNIBS_COILS_DIR")
    if env:
        return Path(env)
    spec = importlib.util.find_spec("cunibs")
    if spec is not None and spec.submodule_search_locations:
        return Path(list(spec.submodule_search_locations)[0]) / "coils"
    return Path("coils")


COILS_ROOT = _root()

FILES = {
    "DEYMED_50BF": "Deymed_50BF.h5",
    "DEYMED_70BF": "Deymed_70BF.h5",
    "DEYMED_120BFV": "Deymed_120BFV.h5",
    "MAGMORE_PMD25_DECENTRAL": "MagMore_PMD25-decentral.h5",
    "MAGMORE_PMD45_EEG": "MagMore_PMD45-EEG.h5",
    "MAGMORE_PMD70": "MagMore_PMD70.h5",
    "MAGSTIM_D70": "MagStim_D70.h5",
    "MAGSTIM_DCC": "MagStim_DCC.h5",
    "MAGVENTURE_C_100": "MagVenture_C-100.h5",
    "MAGVENTURE_C_B60": "MagVenture_C-B60.h5",
    "MAGVENTURE_C_B70": "MagVenture_C-B70.h5",
    "MAGVENTURE_COOL_B35": "MagVenture_Cool-B35.h5",
    "MAGVENTURE_COOL_B65": "MagVenture_Cool-B65.h5",
    "MAGVENTURE_COOL_B70": "MagVenture_Cool-B70.h5",
    "MAGVENTURE_COOL_D_B80": "MagVenture_Cool-D-B80.h5",
    "MAGVENTURE_MC_125": "MagVenture_MC-125.h5",
    "MAGVENTURE_MC_125_NEW": "MagVenture_MC-125_new.h5",
    "MAGVENTURE_MC_B70": "MagVenture_MC-B70.h5",
    "MAGVENTURE_MCF_75": "MagVenture_MCF-75.h5",
    "MAGVENTURE_MCF_B65": "MagVenture_MCF-B65.h5",
    "MAGVENTURE_MCF_B65_NEW": "MagVenture_MCF-B65_new.h5",
    "MAGVENTURE_MC_B65_HO8": "MagVenture_MC_B65_HO8.h5",
    "MAGVENTURE_MMC_140_II": "MagVenture_MMC-140-II.h5",
    "MAGVENTURE_MRI_B91": "MagVenture_MRI-B91.h5",
    "MAGVENTURE_MST_TWIN": "MagVenture_MST_Twin.h5",
}

globals().update({name: COILS_ROOT / fname for name, fname in FILES.items()})
