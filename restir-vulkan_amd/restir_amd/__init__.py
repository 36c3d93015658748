"""restir_amd — MI355X (gfx950) ReSTIR hot path of lukedan/ReSTIR-Vulkan behind a C ABI.

The compute lives in librestir_amd.so (HIP kernels + host builders); this package binds it with
ctypes, generates seeded synthetic scenes and drives frames in the reference's pass order.
"""
from . import abi, scenes  # noqa: F401
from .abi import (Context, RestirError, build_aabb_tree, build_alias_table,  # noqa: F401
                  load_library, make_uniforms, projection_view)

__all__ = ["abi", "scenes", "Context", "RestirError", "build_aabb_tree", "build_alias_table",
           "load_library", "make_uniforms", "projection_view"]
