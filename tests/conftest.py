import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_feat():
    """The package directory has a dot in its name: import it by path."""
    if "feat3amd" in sys.modules:
        return sys.modules["feat3amd"]
    spec = importlib.util.spec_from_file_location(
        "feat3amd", os.path.join(ROOT, "feat3.doxytest_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["feat3amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def feat():
    return load_feat()


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.pyoracle import Reference, LIB_REF
    if not os.path.exists(LIB_REF):
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    return Reference()


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "reference_vectors.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="session")
def handle(feat, torch_gpu):
    # on torch's stream, so tensor set-up and library kernels are ordered
    h = feat.Handle(0, stream=torch_gpu.cuda.current_stream().cuda_stream)
    yield h
    h.close()
