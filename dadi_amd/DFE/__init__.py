"""Selection models and batched DFE cache generation on the device."""
from . import DemogSelModels  # noqa: F401
