"""MI355X-native trace path for MeerCADRaytracer scenes.

Public surface (mirroring the reference's src/index.ts for the trace path):
  raytrace, raytrace_compiled, trace_settings       src/raytrace.ts
  Tracer (render_tile), camera rays                 src/tracer.ts (on the GPU through libmcr.so)
  ProgressiveRenderer, SceneSync, make_tiles        src/progressive.ts
  create_tile_handler, executors                    src/protocol.ts, src/executors.ts
  compile_scene, compile_camera                     src/compile.ts, src/textures.ts (TextureTable)
The GPU work runs in libmcr.so (include/mcr.h); see meercadraytracer_amd.native.
"""
from . import abi, three  # noqa: F401
from .compile import compile_camera, compile_geometry, compile_scene  # noqa: F401
from .raytrace import raytrace, raytrace_compiled, trace_settings  # noqa: F401
from .types import (CompiledCamera, CompiledGeometry, CompiledLight, CompiledMaterial,  # noqa: F401
                    CompiledScene, ShadingPayload, Tile, TraceSettings)
