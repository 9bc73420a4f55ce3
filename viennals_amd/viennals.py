"""Drop-in counterpart of the reference's top-level Python module viennals
(python/__init__.py:56-99) for the advection path: setDimension /
getDimension, and every class resolved in the module of the current
dimension (viennals_amd.d2 by default, as PROXY_DIM = 2 there).
`import viennals_amd.viennals as viennals` runs reference scripts such as
examples/VoidEtching/VoidEtching.py up to their mesh output."""
from . import d2, d3
from . import BooleanOperationEnum, BoundaryConditionEnum, SpatialSchemeEnum, TemporalSchemeEnum  # noqa: F401

PROXY_DIM = 2


def setDimension(d):
    global PROXY_DIM
    if d in (2, 3):
        PROXY_DIM = d
    else:
        raise ValueError("Dimension must be 2 or 3.")


def getDimension():
    return PROXY_DIM


def __getattr__(name):
    m = d2 if PROXY_DIM == 2 else d3
    try:
        return getattr(m, name)
    except AttributeError:
        raise AttributeError(f"module {__name__!r} has no attribute {name!r}") from None
