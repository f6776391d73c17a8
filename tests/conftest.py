import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a visible GPU")
    return torch


@pytest.fixture
def term_split():
    """Set the library's term-split switch for one test; restored after."""
    from scrs_amd.codec import set_term_split

    prev = []

    def set_(enable):
        prev.append(set_term_split(enable))

    yield set_
    if prev:
        set_term_split(prev[0])


@pytest.fixture
def afft_targeted():
    """Set the additive-FFT targeted-path threshold for one test; restored
    after."""
    from scrs_amd.codec import afft_set_targeted_max

    prev = []

    def set_(max_missing):
        prev.append(afft_set_targeted_max(max_missing))

    yield set_
    if prev:
        afft_set_targeted_max(prev[0])
