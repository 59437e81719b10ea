import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O._load()
    return O


@pytest.fixture(scope="session")
def gpu():
    """Torch device handle; fails loudly if the HIP library or the GPU is missing."""
    import torch
    import lifted_amd
    from lifted_amd import _lib
    _lib.lib()
    if not torch.cuda.is_available():
        pytest.fail("GPU test requested but no ROCm device is visible")
    return torch.device("cuda:0")


def pytest_terminal_summary(terminalreporter):
    """Measured parity of every GPU comparison (printed under -q too)."""
    try:
        import helpers
    except ImportError:
        return
    if not helpers.PARITY:
        return
    by = {}
    for test, what, same, err in helpers.PARITY:
        key = test.split("::")[0]
        n, ex, worst, wwhat = by.get(key, (0, 0, 0.0, ""))
        if err > worst:
            worst, wwhat = err, f"{test.split('::')[-1]}: {what}"
        by[key] = (n + 1, ex + (1 if same else 0), worst, wwhat)
    tr = terminalreporter
    tr.write_sep("-", "parity against the oracle (per-band ulp of each band's max |ref|)")
    for key, (n, ex, worst, wwhat) in sorted(by.items()):
        line = f"{key}: {n} comparisons, {ex} bit-exact, worst {worst:.3g} ulp"
        if worst > 0:
            line += f" ({wwhat})"
        tr.write_line(line)
