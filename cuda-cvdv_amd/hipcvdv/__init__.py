"""hipcvdv - MI355X-native hybrid CV-DV state-vector engine (drop-in for CudaCvdv)."""

from .batch import HipCvdvBatch
from .hipCvdv import HipCvdv
from .separable import SeparableState

__all__ = ["HipCvdv", "HipCvdvBatch", "SeparableState"]
