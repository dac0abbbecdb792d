"""MI355X lookahead approximate convex decomposition.

Python mirror of the reference's public surface (src/mod.ts:2-10):
``decompose(mesh, **options)``, ``Decomposer``, ``Part``, ``GpuTree``,
``merge_hulls`` and the mesh helpers, over the C ABI of ``libcuacd.so``
(hand-written HIP kernels for gfx950). There is no CPU fallback: without the
built library every entry point raises.
"""
from ._lib import load_product, library_path, CuacdError, AbortError  # noqa: F401
from .mesh import Mesh, make_box, unit_cube, l_shape, merge_meshes, signed_volume, volume, is_closed, bbox  # noqa: F401
from .mesh import normalize, plane_cut, concave_edge_planes, parse_stl, generate  # noqa: F401
from .mesh import parse_obj, to_obj, clean_hull_faces, compact_hull, hull_eps  # noqa: F401
from .gputree import GpuTree, TreeConfig, DEFAULT_TREE_CONFIG, HD_MAX_PAIRS  # noqa: F401
from .decompose import decompose, decompose_batch, merge_hulls, Decomposer, Part, DecomposeOptions, DEFAULT_OPTIONS, K_RV, rv_cost, part_cost, AbortFlag  # noqa: F401
from .decompose import release_cached_memory  # noqa: F401
