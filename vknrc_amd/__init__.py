"""MI355X-native neural-radiance-cache hot path (encode -> fused 64-wide MLP ->
relative-L2 -> backward -> Adam) behind the C ABI of include/nrc.h."""
from .nrc import *  # noqa: F401,F403
from .nrc import NRCConfig, NRCError, NRCState, load_library  # noqa: F401
