import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is present; the HIP library must then load (no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zodipy_amd import _lib

    _lib.load()
    return torch.device("cuda")


def pytest_report_header(config):
    return f"zodipy_amd root: {ROOT}; HIP lib: {os.environ.get('ZODI_HIP_LIB', 'in-tree')}"


@pytest.fixture(scope="session")
def cloud_ring_model():
    """Name of a registered two-component Kelsall model (DIRBE's cloud and ring).

    Its components form no registry layout, so the forward and adjoint kernels
    run it through their generic paths.
    """
    import dataclasses

    from zodipy_amd.component import ComponentLabel
    from zodipy_amd.zodiacal_light_model import model_registry

    name = "test-cloud-ring"
    try:
        model_registry.get_model(name)
    except ValueError:
        d = model_registry.get_model("dirbe")
        keep = [ComponentLabel.CLOUD, ComponentLabel.RING]
        model_registry.register_model(name, dataclasses.replace(
            d, comps={k: d.comps[k] for k in keep},
            emissivities={k: d.emissivities[k] for k in keep},
            albedos={k: d.albedos[k] for k in keep}))
    return name


@pytest.fixture(scope="session")
def planck18_style():
    """Name of the registered six-component Planck-2018-style Kelsall model.

    The registry's planck18 has four components (cloud and three bands), so the ring
    and the feature take their emissivities from planck13; geometry, T_0 and delta
    are DIRBE's.
    """
    from zodipy_amd import component_params as cp
    from zodipy_amd import source_params as sp
    from zodipy_amd.component import ComponentLabel as L
    from zodipy_amd.zodiacal_light_model import Kelsall, model_registry

    name = "planck18-style-6comp"
    try:
        model_registry.get_model(name)
    except ValueError:
        emiss = {**sp.EMISSIVITY_PLANCK_18, L.RING: sp.EMISSIVITY_PLANCK_13[L.RING],
                 L.FEATURE: sp.EMISSIVITY_PLANCK_13[L.FEATURE]}
        model_registry.register_model(name, Kelsall(
            comps=cp.DIRBE, spectrum=sp.SPECTRUM_PLANCK, emissivities=emiss,
            T_0=sp.T_0_DIRBE, delta=sp.DELTA_DIRBE))
    return name
