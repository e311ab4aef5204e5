"""sol_amd: the path-tracing hot path of sol (madmann91/sol) on MI355X (gfx950).

Product libraries (built in-tree by `make -C sol_amd`):
  libsolhip.so    hand-written HIP kernels + the solhip_* C ABI (include/solhip.h)
  libsol_host.so  C++20 scene model, BVH builder and flattening + the sol_* C ABI (include/sol_host.h)
"""
from . import _ffi  # noqa: F401  (raises ImportError if the libraries are not built)
from .gpu import Context, GpuPathTracer, GpuScene, PathTracerConfig, new_image  # noqa: F401
from .scene import (Scene, camera_basis, default_bvh_config, default_sphere_field_params, gen_rays,  # noqa: F401
                    save_image)
