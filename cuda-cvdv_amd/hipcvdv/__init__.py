"""hipcvdv - MI355X-native hybrid CV-DV state-vector engine (drop-in for CudaCvdv)."""

from .batch import HipCvdvBatch
from .hipCvdv import HipCvdv, HipCvdvGraph
from .separable import SeparableState

__all__ = ["HipCvdv", "HipCvdvBatch", "HipCvdvGraph", "SeparableState"]
