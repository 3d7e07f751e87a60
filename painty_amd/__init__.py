"""MI355X (gfx950) implementation of painty's KM composite + imprint-stamp path.

The compute lives in libpainty_hip.so (C ABI: include/painty_hip.h); the C++
drop-in classes are under painty_amd/include/painty/, and painty_amd.hip is
the Python mirror of the same interface.  Importing the package does not load
the library; the first compute call does, and raises if it is missing.
"""
__version__ = "0.1.0"
