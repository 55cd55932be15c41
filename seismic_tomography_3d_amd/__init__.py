"""MI355X-native first-arrival traveltime tomography.

Python access to the C ABI (include/st_api.h) implemented by hand-written
HIP kernels for gfx950 in libst_hip.so.  The C++ mirror of the reference's
Eikonal / Tomography classes lives in host/ and links the same library.
"""
from .api import Context, StError  # noqa: F401
