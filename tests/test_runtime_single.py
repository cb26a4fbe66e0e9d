"""One HIP runtime per process: the library's package loaded first and PyTorch imported afterwards (the order of
the GPU tests) map one libamdhip64 and one libhsa-runtime64, as does the other order. Two copies of the HSA
runtime in one process manage the same GPU side by side; the GPU tests that imported torch after the library
ended in GPU memory faults while the wheel's copies were mapped beside the installation's."""

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

_PROBE = r"""
import sys
sys.path.insert(0, {pkg!r})
{first}
{second}
files = sorted({{l.split()[-1] for l in open('/proc/self/maps')
                 if 'libamdhip64' in l or 'libhsa-runtime64' in l}})
print(len([f for f in files if 'libamdhip64' in f]), len([f for f in files if 'libhsa-runtime64' in f]))
"""


@pytest.mark.parametrize("torch_first", [False, True])
def test_one_hip_runtime_with_torch(torch_first):
    pytest.importorskip("torch")
    lib = ROOT / "power-grid-model_amd" / "power_grid_model_amd" / "_lib" / "libpower_grid_model_c.so"
    if not lib.exists():
        pytest.skip("library not built")
    ours, theirs = "import power_grid_model_amd", "import torch"
    code = _PROBE.format(pkg=str(ROOT / "power-grid-model_amd"), first=theirs if torch_first else ours,
                         second=ours if torch_first else theirs)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split()[-2:] == ["1", "1"], (out.stdout, out.stderr)
    assert "second HIP runtime" not in out.stderr
