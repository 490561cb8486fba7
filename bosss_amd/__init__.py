"""bosss_amd: MI355X (gfx950) DG spatial-operator hot path of BoSSS.

The compute path is the HIP library libbosss_hip.so (C ABI in include/bosss_hip.h);
this package is the host-side mirror of the reference's operator interface.
"""
from .grid import CartesianGrid  # noqa: F401
from .refelement import Tables  # noqa: F401


def __getattr__(name):
    if name in ("DifferentialOperator", "EvaluatorNonLin", "Context"):
        from . import operator
        return getattr(operator, name)
    raise AttributeError(name)
