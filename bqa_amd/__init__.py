"""MI355X-native backend for belief-propagation tensor-network quantum annealing.

Drop-in for the annealing path of LuchnikovI/bqa: the same config format and
``run_qa`` result format, with ``"backend": "mi355x"`` running every phase of
a Trotter step as HIP kernels on one GPU (``libbqa_hip.so``).
"""
__version__ = "0.1.0"

from .core import run_qa  # noqa: E402,F401
