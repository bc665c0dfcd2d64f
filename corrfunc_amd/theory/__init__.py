"""Theory pair counters (mirror of Corrfunc.theory)."""
from .DD import DD
from .DDrppi import DDrppi
from .DDsmu import DDsmu
from .xi import xi
from .wp import wp

__all__ = ("DD", "DDrppi", "DDsmu", "xi", "wp")
