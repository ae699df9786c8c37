"""MI355X-native CNS_NoSpec time step.

The compute lives in ``libcns_hip.so`` (hand-written HIP for gfx950 behind the
C ABI of ``include/cns.h``).  This package is its host side:

* :mod:`.layout`  -- the ghosted, padded SoA array layout of the ABI;
* :mod:`.state`   -- seeded initial states;
* :mod:`.ops`     -- the reference's operator interface (fill_boundary,
  ctoprim, hypterm, diffterm, new_advance) over the ABI;
* :mod:`.refio`   -- the reference's text file formats;
* :mod:`.batch`   -- many boxes per launch, hipGraph steps, FP32-flux mode;
* :mod:`.serve`   -- the continuous-batching request server;
* :mod:`.build`   -- builds the library in-tree with hipcc.
"""
from .layout import NC, NG, NQ, BoxLayout  # noqa: F401

__all__ = ["NC", "NG", "NQ", "BoxLayout"]
