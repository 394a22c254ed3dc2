"""Python handle on libcpt (the MI355X path tracer behind include/cpt.h).

    import sys; sys.path.insert(0, "<repo>/cpp-pt_amd")
    import cpt
    scene = cpt.scenes.box_room()
    with cpt.Context() as ctx:
        ctx.upload(scene)
        ctx.render(512, 512, spp_axis=8)
        img = ctx.read_film()
"""
from . import raybatch, scenes
from ._lib import (ABI_FUNCTIONS, CPT_RENDER_ACCUMULATE, CPT_RENDER_COUNT, CPT_RENDER_KERNEL_TIMING,
                   CPT_RENDER_GUIDE_TRAIN, CPT_RENDER_GUIDED, CPT_RENDER_NEE_ONLY, CPT_RENDER_NO_GRAPH, CPT_RENDER_PPG,
                   CPT_RENDER_PPG_STEP, CPT_RENDER_PURE_PT, CPT_RENDER_LIGHT_TRACE, CPT_RENDER_CORRECTED,
                   CPT_RENDER_STREAM, CPT_E_HIP, CPT_E_INVALID, CPT_E_LIMIT, Context, CptError, LIB_PATH, load_library,
                   load_scene_file, reference_tie_rank)

__all__ = ["scenes", "raybatch", "reference_tie_rank", "load_scene_file", "Context", "CptError", "load_library", "LIB_PATH", "ABI_FUNCTIONS",
           "CPT_RENDER_NO_GRAPH", "CPT_RENDER_KERNEL_TIMING", "CPT_RENDER_COUNT", "CPT_RENDER_ACCUMULATE", "CPT_RENDER_STREAM",
           "CPT_RENDER_GUIDED", "CPT_RENDER_GUIDE_TRAIN", "CPT_RENDER_NEE_ONLY", "CPT_RENDER_PURE_PT", "CPT_RENDER_PPG",
           "CPT_RENDER_PPG_STEP", "CPT_RENDER_LIGHT_TRACE", "CPT_RENDER_CORRECTED", "CPT_E_HIP", "CPT_E_INVALID",
           "CPT_E_LIMIT"]
