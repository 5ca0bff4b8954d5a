from . import checkpoint, stepper
from .minsr import MinSR
from .tdvp import TDVP, imagFun, realFun
from .util import ground_state_search, measure

__all__ = ["checkpoint", "TDVP", "MinSR", "stepper", "measure", "ground_state_search", "realFun", "imagFun"]
