"""MI355X-native collision library for inMyRoom_vulkan's collision path.

HIP kernels for gfx950 behind a C ABI (``include/mi355x_collide.h``):
OBB-tree build, OBB-tree traversal with the 15-axis separating-axis test,
Möller triangle-triangle test with contact output, sweep-and-prune broad
phase, contact aggregation into uncollide rays, skinning and refit, glTF mesh ingest.  ``collide.Collider`` is the Python binding; ``scenes`` generates the
synthetic workloads.
"""
from .collide import (CONTACT_DTYPE, ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, MODE_REFERENCE, PAIR_DTYPE, PAIR_RAYS_DTYPE,
                      PAIR_RESPONSE_DTYPE, QUERY_DTYPE, RAY_DTYPE, NODE_DTYPE, CD_NO_MESH, CD_NO_PARENT, CollideError,
                      Collider, GltfDocument, load_library)

__all__ = ["Collider", "CollideError", "GltfDocument", "CD_NO_MESH", "CD_NO_PARENT", "NODE_DTYPE", "load_library", "CONTACT_DTYPE", "QUERY_DTYPE", "ENTITY_DTYPE", "PAIR_DTYPE",
           "RAY_DTYPE", "PAIR_RAYS_DTYPE", "PAIR_RESPONSE_DTYPE", "MODE_PLAIN", "MODE_BATCHED",
           "MODE_REFERENCE"]
