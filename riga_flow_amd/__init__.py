"""MI355X (gfx950) backend for riga-engine's path-simulation and Phoenix-payoff hot path.

The product is ``libriga_hip.so`` (C ABI in include/riga_hip.h); this package
is its Python host mirror of the reference pricing API.
"""
from .types import (ImpliedVolGrid, LocalVolGrid, LocalVolSpec, MarketSnapshot, PhoenixTrade,
                    PriceRequest, PriceResult, RigaError, SimulationConfig, VarianceReduction)

__all__ = ["ImpliedVolGrid", "LocalVolGrid", "LocalVolSpec", "MarketSnapshot", "PhoenixTrade",
           "PriceRequest", "PriceResult", "RigaError", "SimulationConfig", "VarianceReduction"]
