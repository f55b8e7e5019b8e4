"""parallelpopgen_amd -- MI355X (gfx950) implementation of GO_Fish's
per-generation Wright-Fisher update and the Spectrum SFS binning of
ParallelPopGen, behind the reference's C++ API (include/go_fish.cuh,
include/spectrum.h) and a C ABI (include/gofish_capi.h).

Python access goes through the C ABI (`parallelpopgen_amd.capi`); the
compute path is the in-tree HIP library libgofish.so, with no CPU fallback.
"""
from . import capi, scenarios  # noqa: F401
from .capi import run, site_frequency_spectrum, population_frequency_histogram, load  # noqa: F401
