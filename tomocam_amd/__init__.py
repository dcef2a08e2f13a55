"""MI355X-native implementation of tomoCAM's MBIR hot path.

Drop-in host interface of the reference's ``tomocam`` package:
``MBIR``, ``recon``, ``radon``, ``radon_adj`` (numpy in / out) and the
``cTomocam`` module (``mbir``, ``radon``, ``backproject``,
``set_num_of_gpus``).  Device-level operators on torch CUDA tensors are in
:mod:`tomocam_amd.gpu`.  All arithmetic runs in ``libtomocam_hip.so``.
"""

from .modeling import MBIR, recon, radon, radon_adj
from . import cTomocam

__version__ = "0.1.0"

__all__ = ["MBIR", "recon", "radon", "radon_adj", "cTomocam", "__version__"]
