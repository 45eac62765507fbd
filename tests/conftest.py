import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cnn-crf-stereo_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    if "cnn_crf_stereo_amd" in sys.modules:
        return sys.modules["cnn_crf_stereo_amd"]
    spec = importlib.util.spec_from_file_location("cnn_crf_stereo_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cnn_crf_stereo_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def handle():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cnn_crf_stereo_amd import capi

    h = capi.Handle()
    yield h
    h.close()
