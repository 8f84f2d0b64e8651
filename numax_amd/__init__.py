"""numax_amd -- MI355X-native kernels for numax.linalg's dense factor-and-solve
path, behind numax's own call signatures (see ``numax_amd.linalg``)."""
from . import random  # noqa: F401  (host-side seeded inputs; no GPU needed)

__version__ = "0.1.0"
