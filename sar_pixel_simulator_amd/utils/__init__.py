from .rounding import round_to_odd
from . import log

__all__ = ['round_to_odd', 'log']
