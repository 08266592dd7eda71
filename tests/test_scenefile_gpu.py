"""A reference-format scene file of the Cornell box (raw meshes, area light,
diffuse BSDFs, lookat pinhole, film asset, renderer node) renders the same
paths as the same scene built directly: per-sample contributions from one
Philox stream agree, and render_file writes the film with its type's suffix."""
import numpy as np
import pytest

from lmx import image, meshfile, render, scenefile, scenes

pytestmark = pytest.mark.gpu


def _cornell_yaml(sc, spp, renderer=None):
    cam = sc["camera"]
    lines = ["lightmetrica:", "  version: 1.1.0", "  assets:", "    type: assets3", "    params:",
             "      film_1:", "        interface: film", "        type: hdr",
             f"        params: {{w: {cam['width']}, h: {cam['height']}, type: radiancehdr}}",
             "      sensor_1:", "        interface: sensor", "        type: pinhole",
             f"        params: {{film: film_1, fov: {cam['fov_deg']}}}"]
    nodes = ["        - id: cam", "          sensor: sensor_1", "          transform:", "            lookat:",
             "              eye: %g %g %g" % cam["eye"], "              center: %g %g %g" % cam["center"],
             "              up: %g %g %g" % cam["up"]]
    light_tris = set()
    for first, count in sc["lights"]:
        light_tris.update(range(first, first + count))
    for mi, m in enumerate(sc["materials"]):
        sel = np.nonzero(sc["tri_mat"] == mi)[0]
        if sel.size == 0:
            continue
        f = sc["tris"][sel]
        used, inv = np.unique(f.reshape(-1), return_inverse=True)
        mesh = meshfile.dump_raw_mesh(sc["verts"][used], inv.reshape(-1, 3))
        lines += [f"      mesh_{mi}:", "        interface: trianglemesh", "        type: raw", "        params:"]
        lines += ["          " + l for l in mesh.splitlines()]
        nodes += [f"        - id: n{mi}", f"          mesh: mesh_{mi}"]
        if m["type"] == scenes.MAT_DIFFUSE:
            lines += [f"      bsdf_{mi}:", "        interface: bsdf", "        type: diffuse",
                      "        params: {R: '%r %r %r'}" % tuple(m["R"])]
            nodes.append(f"          bsdf: bsdf_{mi}")
        if any(m["Le"]):
            lines += [f"      light_{mi}:", "        interface: light", "        type: area",
                      "        params: {Le: '%r %r %r'}" % tuple(m["Le"])]
            nodes.append(f"          light: light_{mi}")
    n = cam["width"] * cam["height"] * spp
    lines += ["  scene:", "    type: scene3", "    params:", "      sensor: cam", "      nodes:"] + nodes
    lines += renderer or ["  renderer:", "    type: pt", "    params:", f"      num_samples: {n}",
                          "      max_num_vertices: 9"]
    return "\n".join(lines) + "\n"


def test_scene_file_renders_the_same_paths(gpu, tmp_path):
    sc = scenes.cornell_box(32, 32)
    text = _cornell_yaml(sc, 16)
    loaded = scenefile.load_scene(text)
    s2 = loaded["scene"]
    assert s2["tris"].shape == sc["tris"].shape and len(s2["lights"]) == 1
    n = 32 * 32 * 16
    a = render.Scene(sc).render_pt(n, max_num_vertices=9, seed=5, sample_out=True)[2].cpu().numpy()
    b = render.Scene(s2).render_pt(n, max_num_vertices=9, seed=5, sample_out=True)[2].cpu().numpy()
    same = (a[:, 5] == b[:, 5]) & np.all(np.abs(a[:, 2:5] - b[:, 2:5]) <= 1e-5 * np.maximum(np.abs(a[:, 2:5]), 1e-3), axis=1)
    assert (a[:, 5] > 0).sum() > 20
    assert same.mean() > 0.999, (~same).sum()
    # the whole command: renderer node, film type, suffix
    path = tmp_path / "cornell.yml"
    path.write_text(text)
    out = scenefile.render_file(str(path), str(tmp_path / "result"), seed=7)
    assert out == str(tmp_path / "result") + ".hdr"
    img = image.load(out)
    ref = render.Scene(sc).render_pt(n, max_num_vertices=9, seed=7, raster_mode=render.RASTER_RANDOM)[0].cpu().numpy()
    # RGBE keeps 8 mantissa bits of the largest channel
    assert img.shape == ref.shape
    np.testing.assert_allclose(img.mean(), ref.mean(), rtol=1e-2)
    m = ref.max(axis=2, keepdims=True)
    assert (np.abs(img - ref) <= m * 2.0 ** -7 + 1e-6).mean() > 0.99


@pytest.mark.parametrize("mode", ["vcm", "bdpm", "bdpt"])
def test_scene_file_vcm_renderer_node(gpu, mode):
    """renderer::vcm's node (renderer_vcm.cpp:558-577): every mode renders
    num_iteration_pass passes of num_eye_trace_samples eye samples."""
    sc = scenes.cornell_box(32, 32)
    node = ["  renderer:", "    type: vcm", "    params:", f"      mode: {mode}", "      num_iteration_pass: 3",
            "      num_eye_trace_samples: 8192", "      num_photon_trace_samples: 5000",
            "      initial_radius: 0.05", "      alpha: 0.5", "      max_num_vertices: 6"]
    loaded = scenefile.load_scene(_cornell_yaml(sc, 1, node))
    film, st = scenefile.render(loaded, seed=11)
    rs = render.Scene(sc)
    kw = dict(max_num_vertices=6, seed=11, raster_mode=render.RASTER_RANDOM)
    if mode == "bdpt":
        ref = rs.render_vcm_bdpt(3 * 8192, **kw)[0]
    else:
        ref = rs.render_vcm(8192, mode=render.VCM_MODE_VCM if mode == "vcm" else render.VCM_MODE_BDPM,
                            num_iteration_pass=3, num_photon_trace_samples=5000, initial_radius=0.05, alpha=0.5,
                            **kw)[0]
    a, b = film.cpu().numpy(), ref.cpu().numpy()
    assert a.mean() > 0
    # the same samples; splat order differs between the two renders
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.max()))


def test_scene_file_pm_renderer_node(gpu):
    """renderer::pm's node (renderer_pm.cpp:61-70) renders as lmx_render_pm
    with the same parameters."""
    sc = scenes.cornell_box(32, 32)
    node = ["  renderer:", "    type: pm", "    params:", "      num_samples: 16384",
            "      num_photon_trace_samples: 20000", "      finalgather: 1", "      radius: 0.05",
            "      max_num_vertices: 7", "      photonmap: kdtree"]
    film, _ = scenefile.render(scenefile.load_scene(_cornell_yaml(sc, 1, node)), seed=13)
    ref = render.Scene(sc).render_pm(16384, num_photon_trace_samples=20000, finalgather=1, radius=0.05,
                                     max_num_vertices=7, seed=13, raster_mode=render.RASTER_RANDOM)[0]
    a, b = film.cpu().numpy(), ref.cpu().numpy()
    assert a.mean() > 0
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.max()))


def test_scene_file_ppm_renderer_node(gpu):
    """renderer::ppm's node (renderer_ppm.cpp:76-89) renders as
    lmx_render_ppm with the same parameters."""
    sc = scenes.cornell_box(32, 32)
    node = ["  renderer:", "    type: ppm", "    params:", "      num_samples: 4096", "      max_num_vertices: 7",
            "      num_iteration_pass: 4", "      num_photon_trace_samples: 20000", "      initial_radius: 0.05",
            "      alpha: 0.6"]
    film, _ = scenefile.render(scenefile.load_scene(_cornell_yaml(sc, 1, node)), seed=17)
    ref = render.Scene(sc).render_ppm(4096, 7, num_iteration_pass=4, num_photon_trace_samples=20000,
                                      initial_radius=0.05, alpha=0.6, seed=17, raster_mode=render.RASTER_RANDOM)[0]
    a, b = film.cpu().numpy(), ref.cpu().numpy()
    assert a.mean() > 0
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.max()))


def test_scene_file_sppm_renderer_node(gpu):
    """renderer::sppm's node (renderer_sppm.cpp:81-96) renders as
    lmx_render_sppm with the same parameters."""
    sc = scenes.cornell_box(32, 32)
    node = ["  renderer:", "    type: sppm", "    params:", "      max_num_vertices: 7",
            "      num_iteration_pass: 3", "      num_photon_trace_samples: 20000", "      initial_radius: 0.05"]
    film, _ = scenefile.render(scenefile.load_scene(_cornell_yaml(sc, 1, node)), seed=19)
    ref = render.Scene(sc).render_sppm(7, num_iteration_pass=3, num_photon_trace_samples=20000, initial_radius=0.05,
                                       seed=19)[0]
    a, b = film.cpu().numpy(), ref.cpu().numpy()
    assert a.mean() > 0
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.max()))
