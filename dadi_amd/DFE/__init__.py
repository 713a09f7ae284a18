"""Selection models and batched DFE cache generation on the device."""
from . import DemogSelModels  # noqa: F401
from . import PDFs  # noqa: F401,E402
from .Cache1D_mod import Cache1D  # noqa: F401,E402
from .Cache2D_mod import Cache2D  # noqa: F401,E402
