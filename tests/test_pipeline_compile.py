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


def _resources(driver, scene, tile_shift):
    rc, txt = compile_offline(driver, scene, 2, tile_shift=tile_shift)
    assert rc == 0, txt
    return {k: int(v) for k, v in (kv.split("=") for kv in txt.split())}


@pytest.mark.parametrize("name,tile_shift,samples", [("grid_textured_blend", 4, 1), ("grid_textured_blend", 5, 1),
                                                      ("phong_msaa4", 4, 4)])
def test_raster_kernel_resources(driver, name, tile_shift, samples):
    """The kernel-descriptor reader against what the raster's LDS must be: per
    pixel colour and depth per sample and one claim word, plus the fragment
    queue (4 entries per pixel), 4 bytes each; registers in granules of 8 and
    the occupancy they and the LDS allow (512 VGPRs per SIMD lane, 160 KB of
    LDS per CU, one-wave workgroups)."""
    r = _resources(driver, _scenes()[name], tile_shift)
    tp = 1 << (2 * tile_shift)
    for k in ("kz_raster", "kz_raster_persistent"):
        assert r[k + "_lds"] == tp * samples * 8 + tp * 4 + tp * 4 * 4
        v = r[k + "_vgprs"]
        assert v % 8 == 0 and 8 <= v <= 512
        assert r[k + "_waves"] == max(1, min(512 // v, 160 * 1024 // r[k + "_lds"] // 4, 8))
    assert r["kz_raster_scratch"] == 0


def test_occupancy_step_raises_only_without_spills(driver):
    """A raster kernel compiled again for one more resident wave keeps it only
    without scratch: the grid pipeline's one-workgroup-per-tile raster (one
    granule above 5 waves at the default hint) runs at 5 waves per SIMD with
    no scratch."""
    r = _resources(driver, _scenes()["grid_textured_blend"], 4)
    assert r["kz_raster_waves"] >= 5 and r["kz_raster_scratch"] == 0
    assert r["kz_raster_persistent_scratch"] == 0
