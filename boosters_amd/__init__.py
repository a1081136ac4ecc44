"""boosters_amd — MI355X (gfx950) histogram GBDT engine with egordm/boosters' API.

The Python surface mirrors the reference bindings
(packages/boosters-python/python/boosters/_boosters_rs.pyi): ``Dataset``,
``GBDTConfig``, ``GBDTModel.train/predict/predict_raw``, ``Objective``,
``Metric``, ``GrowthStrategy``.  All numeric work runs in the native library
``libboosters_amd.so`` (HIP kernels for gfx950); importing the model classes
does not need a GPU, but training and prediction fail loudly without one.
"""
from .model import (  # noqa: F401
    Context,
    Dataset,
    GBDTConfig,
    GBDTModel,
    GrowthStrategy,
    Metric,
    Objective,
    Schedule,
    Server,
)

__all__ = [
    "Context",
    "Dataset",
    "GBDTConfig",
    "GBDTModel",
    "GrowthStrategy",
    "Metric",
    "Objective",
    "Schedule",
    "Server",
]
