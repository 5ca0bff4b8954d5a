from .base import Operator
from .branch_free import (BranchFreeOperator, Id, LocalOp, OpStr, Sm, Sp, Sx, Sy, Sz, scal_opstr)

__all__ = ["Operator", "BranchFreeOperator", "Id", "LocalOp", "OpStr", "Sx", "Sy", "Sz", "Sp", "Sm", "scal_opstr"]
