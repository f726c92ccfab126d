"""renderlab_amd: MI355X-native software BVH ray query for renderlab's non-RT path.

The compute path is librq.so (C++ host code and hand-written gfx950 HIP
kernels behind the C ABI in include/rq.h); ``renderlab_amd.rq`` is its ctypes
mirror and fails loudly on import when the library is missing. Build with
``python -m renderlab_amd.build``.
"""
_EXPORTS = ("Scene", "DeviceBuffer", "Stream", "Event", "build_bvh", "build_bvh_sah", "refit_bvh", "instance_bvh",
            "gi_closest", "gi_visible", "gi_visible_within", "RAY_DTYPE", "HIT_DTYPE")


def __getattr__(name):
    # Loaded on first use so that `renderlab_amd.build` can run before librq.so exists.
    if name in _EXPORTS:
        from . import rq
        return getattr(rq, name)
    raise AttributeError(name)
