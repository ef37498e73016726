"""Pipeline compilation without a device: SPIR-V -> IR -> HIP -> gfx950.

kazan_compile_pipeline_spirv runs what vkCreateGraphicsPipelines runs
(api_impl.rs:4018-4032 -> GraphicsPipeline::create, pipeline.rs:165-183 in
the reference, where it is a stub) for every shader pair and fixed-function
combination the scene workloads use; the code objects land in the on-disk
cache that the device path loads from.
"""
import pytest

from kazan_amd.scenes import generators as g
from kazan_amd.scenes.renderer import compile_offline


def _scenes():
    out = {
        "pos4_color4": g.random_triangles(8, 32, 32),
        "grid_textured_blend": g.jittered_grid(4, 4, 32, 32, 16),
        "mvp_color4": g.zipf_draws(4, 64, 32, 32),
        "cluster_blend": g.gaussian_cluster(64, 32, 32),
        "phong_msaa4": g.stress_4k(4, 64, 32, 32, 4),
    }
    s = g.random_triangles(8, 32, 32)
    s.cull_mode, s.front_face, s.depth_op = 2, 1, 4
    out["cull_back_cw_greater"] = s
    return out


@pytest.mark.parametrize("name", sorted(_scenes()))
def test_compile_code_object(driver, name):
    s = _scenes()[name]
    rc, txt = compile_offline(driver, s, 2)
    assert rc == 0, txt
    fields = dict(kv.split("=") for kv in txt.split())
    assert int(fields["code_object_bytes"]) > 10000


def test_generated_source_has_stage_functions(driver):
    rc, src = compile_offline(driver, g.jittered_grid(4, 4, 32, 32, 16), 0)
    assert rc == 0, src
    for sym in ("kz_vs_main", "kz_fs_main", "kz_vertex", "kz_fragment"):
        assert sym in src


def test_full_source_bakes_fixed_function_state(driver):
    s = g.stress_4k(4, 64, 32, 32, 4)
    rc, src = compile_offline(driver, s, 1, tile_shift=5)
    assert rc == 0, src
    assert "#define KZ_SAMPLES 4" in src and "#define KZ_TILE_SHIFT 5" in src
    assert "kz_raster_persistent" in src and "kz_bin_count" in src


def test_second_compile_hits_the_cache(driver):
    s = g.random_triangles(8, 32, 32)
    compile_offline(driver, s, 2)
    rc, txt = compile_offline(driver, s, 2)
    assert rc == 0 and "cache_hit=1" in txt
