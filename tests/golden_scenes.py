"""The seeded scenes whose SoftGLRender outputs are stored as golden fixtures."""
from softglrender_amd import scenes

GOLDEN = {
    "c1": lambda: scenes.scene_c1(),
    "c1_msaa": lambda: scenes.scene_c1(msaa=True, read_msaa=True),
    "c2_small": lambda: scenes.scene_c2(480, 270, freq=40, tex_size=256, shadow_size=512, read_shadow=True,
                                        read_samples=True),
}
# image-based lighting precompute (irradiance, prefiltered levels), skybox (cube and equirectangular) and FXAA
GOLDEN["ibl_cube"] = lambda: scenes.scene_ibl()
GOLDEN["ibl_equirect_fxaa"] = lambda: scenes.scene_ibl(equirect=True, fxaa=True)
# points, lines and triangles in point and line mode, single-sampled and 4x
GOLDEN["points_lines"] = lambda: scenes.scene_points_lines()
GOLDEN["points_lines_msaa"] = lambda: scenes.scene_points_lines(msaa=True, read_samples=True)
for _v in scenes.VARIANTS:
    GOLDEN[f"v_{_v}"] = (lambda v=_v: scenes.scene_variant(v))
