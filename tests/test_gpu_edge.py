"""Edge cases of the frame entry point on the GPU: empty input, one particle,
all particles in one cell, particles outside the grid or NaN, output capacity
overflow.  Each gives a defined status and sane outputs (the reference exits
or indexes out of bounds for several of them)."""
import numpy as np
import pytest

import oracle
from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor, stats_dict
from tests import parity

pytestmark = pytest.mark.gpu


def frame_from(pos, spacing=0.01, grid_res=32, iso=-1e-4):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    seed_pos = pos if pos.shape[0] else np.zeros((1, 3), np.float32)
    fr = ds._finish("custom", seed_pos, spacing, ("cubic", grid_res), iso, "zb05", 0)
    fr.pos = pos
    fr.dens = np.full(pos.shape[0], 1000.0, np.float32)
    return fr


@pytest.mark.parametrize("method", ["zb05", "akinci12"])
def test_empty_input(torch_cuda, method):
    print(f"\n[gpu case] empty {method}", flush=True)
    fr = frame_from(np.zeros((0, 3)))
    rec = Reconstructor(method)
    res = rec.run(fr)
    st = res.stats
    assert st["num_particles"] == 0 and st["num_band_vertices"] == 0
    assert st["num_mesh_vertices"] == 0 and st["overflow"] == 0
    assert float(res.phi.abs().max()) == 0.0


@pytest.mark.parametrize("method", ["zb05", "akinci12"])
def test_single_particle(torch_cuda, method):
    print(f"\n[gpu case] single particle {method}", flush=True)
    fr = frame_from([[0.0123, 0.0456, 0.0311]], grid_res=40)
    rec = Reconstructor(method)
    res = rec.run(fr)
    ref = parity.Stage(fr, method)
    st = res.stats
    assert st["num_surface_particles"] == 1
    assert st["num_band_vertices"] == ref.index.shape[0] > 0
    phi = res.phi.cpu().numpy()
    bad, _ = parity.field_ok(phi, ref.phi, float(fr.h), ref.wsum, ref.index)
    assert bad == 0


def test_all_particles_in_one_cell(torch_cuda):
    print("\n[gpu case] one cell", flush=True)
    rng = ds.SplitMix64(3)
    pts = 0.05 + 0.01 * rng.uniform(3 * 200).reshape(-1, 3)
    fr = frame_from(pts, spacing=0.01, grid_res=48)
    rec = Reconstructor("zb05")
    res = rec.run(fr)
    ref = parity.Stage(fr, "zb05")
    st = res.stats
    assert st["num_surface_particles"] == int(ref.flags.sum())
    assert st["num_band_vertices"] == ref.index.shape[0]
    bad, _ = parity.field_ok(res.phi.cpu().numpy(), ref.phi, float(fr.h), ref.wsum, ref.index)
    assert bad == 0


@pytest.mark.parametrize("method", ["zb05", "akinci12"])
def test_out_of_grid_and_nan(torch_cuda, method):
    print(f"\n[gpu case] out of grid / NaN {method}", flush=True)
    fr = ds.jittered_lattice(n_side=12, grid_res=48)
    pos = fr.pos.copy()
    pos[5] = [50.0, 0.0, 0.0]
    pos[77] = [np.nan, 0.1, 0.1]
    pos[300] = [-3.0, -3.0, -3.0]
    fr.pos = pos
    rec = Reconstructor(method)
    res = rec.run(fr)
    st = res.stats
    assert st["num_out_of_grid"] == 3
    keys = rec.buffer("cell_keys", np.uint32, fr.n)
    assert np.all(keys[-3:] == fr.sp.count) and np.all(keys[:-3] < fr.sp.count)
    hs = oracle.hash_sort(fr.pos, fr.dens, fr.sp)
    assert np.array_equal(keys, hs["keys"])
    assert np.array_equal(rec.buffer("permutation", np.uint32, fr.n), hs["perm"])


def test_capacity_overflow(torch_cuda):
    print("\n[gpu case] capacity overflow", flush=True)
    torch = torch_cuda
    fr = ds.jittered_lattice(n_side=14, grid_res=56)
    rec = Reconstructor("zb05")
    full = rec.run(fr)
    T = full.stats["num_mesh_vertices"]
    ref_pos = full.tri_pos[:T].cpu().numpy().copy()
    assert T > 60
    cap = 60
    pos = torch.from_numpy(fr.pos).cuda()
    dens = torch.from_numpy(fr.dens).cuda()
    rec._tri.fill_(-7.0)
    rec.launch(fr, pos, dens, cap)
    torch.cuda.synchronize()
    st = stats_dict(rec._stats)
    assert st["overflow"] == 1 and st["num_mesh_vertices"] == T
    got = rec._tri[:cap].cpu().numpy()
    written = got[:, 0] != -7.0
    assert np.array_equal(got[written], ref_pos[:cap][written])
    assert np.all(rec._tri[cap:].cpu().numpy() == -7.0)
