"""MI355X-native SU(3) gauge-link hot path for Gaugefields.jl (host-side mirror)."""
from . import generators  # noqa: F401
