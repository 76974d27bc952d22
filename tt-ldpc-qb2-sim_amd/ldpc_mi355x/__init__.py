"""MI355X-native Monte Carlo LDPC decoding for the CCSDS AR4JA codes."""
from . import codes  # noqa: F401
