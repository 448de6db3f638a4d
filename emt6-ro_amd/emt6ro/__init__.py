"""MI355X-native EMT6/Ro tumour-spheroid simulator.

Drop-in for the reference's ``emt6ro`` Python package: ``emt6ro.simulation``
(python/emt6ro/simulation/__init__.py) and ``emt6ro.ga_model``
(python/emt6ro/ga_model/model.py), backed by hand-written HIP kernels for
gfx950 behind a C ABI (include/emt6ro_amd.h).
"""
