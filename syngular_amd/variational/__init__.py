"""DMRG environment blocks on the device (syngular.variational)."""
from .dmrg import DMRG, left_blocks, restore, right_blocks

__all__ = ["DMRG", "left_blocks", "right_blocks", "restore"]
