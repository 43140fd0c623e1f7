"""md_tian2_amd: MI355X-native EMT / propagator / RPMD hot path of md_tian2.

Native parts: ``libmdt_gpu.so`` (HIP kernels for gfx950 + the C-ABI of
include/mdt_gpu.h) and the Fortran host in ``fortran/``.  The Python modules
are the ctypes binding (``mdt``), the ``.pes`` reader (``pesfile``) and the
synthetic-geometry generator (``geometry``).
"""
__all__ = ["mdt", "pesfile", "geometry"]
