"""MI355X MLS-MPM engine: Python host mirror of the reference engine surface over the HIP C-ABI."""
from . import abi, scenes  # noqa: F401
from .abi import MpmParams, load_library, preset  # noqa: F401
from .engine import Engine, MpmError  # noqa: F401
