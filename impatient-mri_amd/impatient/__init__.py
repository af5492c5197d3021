"""Python host side of the MI355X IMPATIENT engine.

``_lib`` binds the C ABI of libimpatient.so; ``ops`` wraps the operators for
numpy arrays; ``files`` reads and writes the reference data formats;
``synth`` makes the synthetic reconstruction problems.
"""
from ._lib import (Context, DeviceArray, ImpError, load, LIB_PATH, HEADER_PATH,  # noqa: F401
                   IMP_CG_BEST_ITERATE, IMP_CG_TRUE_RESIDUAL, IMP_CG_USE_X0, IMP_CG_PRECOND,
                   IMP_CG_FALLBACK_ZERO, IMP_CG_GRAPH,
                   IMP_Q_CUMULATIVE, IMP_Q_DIRECT)
