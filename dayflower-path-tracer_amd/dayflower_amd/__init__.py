"""MI355X renderer for Dayflower's RendererKernel hot path.

``scene``    - compiled scenes (flat CompiledScene arrays) and their builders.
``assets``   - scene construction, Wavefront OBJ models and image textures.
``renderer`` - HIPRendererKernel, the renderer-kernel interface over the C ABI.
"""

from . import _ffi  # noqa: F401
from .assets import SceneBuilder, default_camera, image_rgb, material_showcase_scene, transform  # noqa: F401
from .renderer import (  # noqa: F401
    HIPBatchRenderer,
    HIPError,
    HIPRendererKernel,
    MATH_EXACT,
    MATH_FAST,
    PT_AUTO,
    PT_MEGAKERNEL,
    PT_WAVEFRONT,
    TILE_ROWS,
    TILE_SQUARE,
    TAIL_BOUNCE_DEFAULT,
    check_reciprocal,
    check_terrain_march,
    device_count,
    measure_roof,
)
from .scene import (  # noqa: F401
    CompiledScene,
    build_batch_scene,
    build_config,
    camera_setup,
    default_params,
    diff_checker_map,
    diff_maps,
    pixel_seeds,
    read_scene,
    sky_at,
    sky_set,
    sun_direction,
)
