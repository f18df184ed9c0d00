"""MI355X-native element assembly and sparse solve behind OpenCMISS-Iron's API.

``from iron_amd import iron`` gives the iron-compatible Python API; ``fem`` is the
device engine; ``include/mife.h`` documents the C ABI of libmi_fe.so.
"""
__all__ = ["iron", "fem", "basis", "mesh"]
