import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session."""
    if not gpu_available():
        pytest.skip("no GPU")
    from rustcolmap_amd.mvs import MvsContext
    c = MvsContext(0)
    yield c
    c.close()


_SCENES = {}


def scene(name: str):
    """Rendered scene, cached per process (rendered on the CPU so inputs do
    not depend on the device)."""
    if name not in _SCENES:
        from rustcolmap_amd.scene import config_spec, render_scene
        _SCENES[name] = render_scene(config_spec(name), device="cpu")
    return _SCENES[name]


def view_data(sc):
    from rustcolmap_amd.mvs import ViewData
    return [ViewData.from_scene_view(v, name="v%03d" % i, image_id=i + 1) for i, v in enumerate(sc.views)]


def bits(a: np.ndarray) -> np.ndarray:
    """Bit pattern of a float array (for bit-exact comparisons)."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)
