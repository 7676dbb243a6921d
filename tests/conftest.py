import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the FP64 restatements (checkers only)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ldpc_gpu():
    import lizard_amd
    dec = lizard_amd.LdpcDecoder()
    yield dec
    dec.close()


@pytest.fixture(scope="session")
def decoder_gpu():
    import lizard_amd
    dec = lizard_amd.Decoder(0)
    yield dec
    dec.close()
