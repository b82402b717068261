"""The reference's kernel-loader interface (QuICT/ops/utils/algorithm_loader.py) over the C ABI."""
from .linalg_loader import LinAlgLoader  # noqa: F401
