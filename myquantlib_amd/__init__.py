"""MI355X-native Monte Carlo engine for the ASVJ model family (Heston,
Bates, Bates-Kou, double-factor variants) behind the C-ABI of
include/myql_capi.h. `pricing.Engine` drives libmyql_hip.so from Python."""
from . import capi  # noqa: F401
