"""On-sky (mock catalog) pair counters (mirror of Corrfunc.mocks: DDrppi_mocks, DDsmu_mocks)."""
from .DDrppi_mocks import DDrppi_mocks
from .DDsmu_mocks import DDsmu_mocks

__all__ = ("DDrppi_mocks", "DDsmu_mocks")
