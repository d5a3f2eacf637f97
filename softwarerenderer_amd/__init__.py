"""MI355X (gfx950) implementation of trenki2/SoftwareRenderer's draw path.

The host API mirrors the reference's Rasterizer / VertexProcessor over the C ABI declared in
include/swr_hip.h; the kernels live in csrc/ and build into libswr_hip.so in-tree.
"""
from .api import (CullMode, DrawMode, RasterMode, Rasterizer, VertexProcessor,  # noqa: F401
                  PixelShaderBase, VertexShaderBase, NullPixelShader, DummyVertexShader,
                  DepthCheckerShader, RGBShader, CoverageShader, HashShader, LogShader,
                  MvpVertexShader, PosRGBVertexShader, RawVertexShader, RASTERIZER_VERTEX,
                  VisibilityShader, Batch)
