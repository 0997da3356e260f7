"""mdcore: host side of the MI355X molecular-dynamics core.

The engine itself is ``libmdcore.so`` (hand-written HIP for gfx950 behind a
C ABI, declared in ``include/mdcore.h``).  This package mirrors the upstream
project's ``src/engine`` API over it: ``OpenMMSystem``/``OpenMMState`` in,
``HipEngine`` (the ``ProtocolEngine`` interface) and the protocol writers out.
"""
from .system import (OpenMMSystem, parseSystem, serialiseSystem, UnsupportedForce,  # noqa: F401
                     XmlError, reduceBox, isReducedForm, boxVolume, perpendicularWidths,
                     boxVectorsForShape)
from .state import OpenMMState, parseState, serialiseState  # noqa: F401
