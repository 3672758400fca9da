"""Shared test setup: import paths, the `gpu` marker, and GPU gating.

GPU tests are marked `gpu` and need a visible HIP device plus the built
library; CPU tests (everything else) exercise the oracle, the host builder and
flattener, and the C ABI's host entry points.
"""
import os
import sys

import pytest

try:  # import torch first so the process's one HIP runtime is torch's
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("illum-rt_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and runs the gfx950 kernels")


@pytest.fixture(scope="session")
def torch_cuda():
    """torch with a visible HIP device (imported before the library)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch


@pytest.fixture(scope="session")
def hip_ctx(torch_cuda):
    import hiprt
    ctx = hiprt.Context(0)
    yield ctx
    torch_cuda.cuda.synchronize()
    ctx.close()
