import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# torch ships its own HIP runtime with the same SONAME as /opt/rocm's; the one
# loaded first serves both libgalax_fmm.so and torch. Import torch before the
# library so that tests mixing the two share torch's runtime.
try:
    import torch  # noqa: F401
except Exception:  # torch is optional for the CPU suite
    torch = None
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")
