"""blazert_amd: MI355X-native closest-hit ray queries with blazeRT's scene API.

Host mirror of blazeRT (cstatz/blazert): ``Scene`` (add_mesh / add_spheres / add_planes / add_cylinders,
commit), ``Ray``, ``RayHit``, ``intersect1``, plus the batched ``Scene.intersect`` over device-resident rays.
``BVH`` builds the reference's tree over a user-defined collection, traversed by the caller's device
intersector (include/blazert_amd/user_geometry.h). The traversal and intersection run in hand-written HIP kernels
(libblazert_amd.so, C ABI in include/blazert_amd.h).
"""
from .capi import (FLAG_ANY_HIT, FLAG_HITS_INOUT, GEOM_CYLINDERS, GEOM_PLANES, GEOM_SPHERES, GEOM_TRIANGLES,
                   INVALID_ID, PRECISION_FP32_TRAV_FP64_HIT, PRECISION_FP64, BlazertError)
from .collection import BVH
from .scene import DBL_MAX, Batch, Ray, RayHit, Scene, intersect1

__all__ = ["Scene", "BVH", "Batch", "Ray", "RayHit", "intersect1", "INVALID_ID", "DBL_MAX", "BlazertError", "PRECISION_FP64",
           "PRECISION_FP32_TRAV_FP64_HIT", "FLAG_ANY_HIT", "FLAG_HITS_INOUT", "GEOM_TRIANGLES", "GEOM_SPHERES",
           "GEOM_PLANES", "GEOM_CYLINDERS"]
