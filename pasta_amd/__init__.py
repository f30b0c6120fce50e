"""MI355X-native (gfx950) model-evaluation path of PASTA.

`import pasta_amd as pasta` gives the reference's Python API for this path
(see pasta_amd.api); the compute runs in libpasta_hip.so.
"""
from .api import *  # noqa: F401,F403
from .api import Star  # noqa: F401
