"""The .vis writer (fsr_vis_write, the reference's saveMiddleDataToVisFile:
GPU/ReconstructorGPUOursZB05.cpp:104-231, ReconstructorGPUAkinci12.cpp:290-413,
ReconstructorGPUOursYu13.cpp:285-431) after a frame of each method: every
section equals the frame's own buffers (sorted particles, occupied hash cells,
band vertices, surface particles, valid cubes; for Yu13 the smoothed and the
involved particles), the band vertices and surface particles equal the
oracle's, and the header carries the grids and radii.  Floats are compared to
the 6 significant digits of the iostream format the file uses."""
import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd import frame_io
from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor
from tests import parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["zb05", "akinci12", "yu13"])
def test_vis_sections(torch_cuda, tmp_path, method):
    print(f"\n[gpu case] vis {method}", flush=True)
    if method == "yu13":
        fr = ds.jittered_lattice(n_side=18, grid_res=72, scale=1.3, iso=0.0, method="yu13",
                                 cell=0.004)
        rec = Reconstructor("yu13", aniso=ds.yu13_params(fr, True))
    else:
        fr = ds.jittered_lattice(n_side=20, grid_res=64)
        rec = Reconstructor(method)
    res = rec.run(fr)
    st = res.stats
    path = str(tmp_path / "f.vis")
    frame_io.write_vis(rec, path, fr, st, "f.obj")
    vis = frame_io.read_vis(path)
    n, C = fr.n, fr.sp.count
    for key, g in (("sp", fr.sp), ("sc", fr.sc)):
        assert vis[key]["res"] == list(g.res)
        np.testing.assert_allclose(vis[key]["min"], g.min, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(vis[key]["cell"], float(g.cell), rtol=1e-5)
    np.testing.assert_allclose(vis["particle_radius"], fr.particle_radius, rtol=1e-5)
    np.testing.assert_allclose(vis["smoothing_radius"], fr.h, rtol=1e-5)
    assert vis["obj"] == "f.obj"
    pos = rec.pos_sorted.cpu().numpy()
    assert vis["particles"].shape == (n, 3)
    np.testing.assert_allclose(vis["particles"], pos, rtol=1e-5, atol=1e-7)
    occ = rec.buffer("occupancy", np.float32, C)
    assert np.array_equal(vis["occupied"], np.nonzero(occ > 0)[0])
    B, Q = st["num_band_vertices"], st["num_valid_cubes"]
    assert np.array_equal(vis["band"], rec.buffer("band_index", np.uint32, B).astype(np.int64))
    flags = rec.buffer("surface_flags", np.uint32, n)
    assert np.array_equal(vis["surface"], np.nonzero(flags == 1)[0])
    valid = rec.buffer("valid_cubes", np.uint32, Q).astype(np.int64)
    assert np.array_equal(vis["valid_cubes"], valid)
    assert np.all(np.diff(valid) > 0)
    if method == "yu13":
        sm = rec.buffer("smooth_pos", np.float32, 4 * n).reshape(n, 4)[:, :3]
        np.testing.assert_allclose(vis["smoothed"], sm, rtol=1e-5, atol=1e-7)
        I = st["num_involved_particles"]
        assert I > 0
        assert np.array_equal(vis["involved"],
                              rec.buffer("involved_index", np.uint32, I).astype(np.int64))
    else:
        assert vis["smoothed"].shape[0] == 0 and vis["involved"].shape[0] == 0
        ref = parity.Stage(fr, method)
        if method == "zb05" or np.array_equal(flags, ref.flags):
            assert np.array_equal(vis["band"], ref.index.astype(np.int64))
            assert np.array_equal(vis["surface"], np.nonzero(ref.flags == 1)[0])
