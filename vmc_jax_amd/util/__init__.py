from . import checkpoint, stepper, symmetries
from .minsr import MinSR
from .tdvp import TDVP, imagFun, realFun
from .util import ground_state_search, measure

__all__ = ["checkpoint", "symmetries", "TDVP", "MinSR", "stepper", "measure", "ground_state_search", "realFun", "imagFun"]
