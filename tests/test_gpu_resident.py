"""Registrations that start from images already in HBM (fw_register_resident): the device forms of
the intensity preprocessing (preprocess.rs:159-232) and of the centre of mass (init.rs:47-86)
against the host code both engines share, and the whole registration against fw_register on the
same f32 volumes.

The device bins f32 intensities through f32 thresholds that reproduce the f64 bins exactly, so
the histogram, its quantiles and the window are those of the host; the window map itself is an
f32 pair computation of the f64 one, equal to it rounded once to f32 up to near-ties (at most one
f32 ulp, on a handful of voxels)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd.backend import RegError, read_composite_h5
from fastwarp_amd.phantom import make_pair
from fastwarp_amd.register import cohort_plan, register, register_resident, write_composite
from tests import cases

pytestmark = pytest.mark.gpu


def f32(a: np.ndarray) -> np.ndarray:
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def volumes():
    rng = np.random.default_rng(21)
    pair = make_pair((40, 44, 36), 2.0, seed=5)
    g = pair.grid
    spiky = f32(pair.fixed[0] * 900.0 + 40.0)
    spiky[rng.integers(0, g.n, 30)] = 9000.0  # outliers that the window clips
    flat = f32(np.where(pair.moving[0] > 0.4, 1.0, 0.0) + rng.normal(0, 1e-3, g.n))
    skew = f32(np.exp(3.0 * pair.fixed[0]) - 7.0)
    return g, {"phantom": f32(pair.fixed[0]), "spiky": spiky, "flat": flat, "skew": skew}


@pytest.mark.parametrize("name", ["phantom", "spiky", "flat", "skew"])
@pytest.mark.parametrize("window", [(0.005, 0.995), (0.0, 1.0), (0.1, 0.8)])
def test_device_preprocess_matches_host(backend, name, window):
    g, vols = volumes()
    v = vols[name]
    want = f32(oracle.preprocess(g, v, *window))
    got = backend.download_image(backend.preprocess_image(backend.upload_image(g, v), *window))
    diff = np.abs(got - want)
    assert diff.max() <= np.spacing(np.float32(1.0)), (name, window, diff.max())
    assert np.count_nonzero(diff) <= 10, (name, window, np.count_nonzero(diff))


def test_device_preprocess_rejects_non_finite(backend):
    g, vols = volumes()
    v = vols["phantom"].copy()
    v[123] = np.nan
    with pytest.raises(RegError, match="non-finite"):
        backend.preprocess_image(backend.upload_image(g, v))


@pytest.mark.parametrize("name", ["phantom", "spiky", "skew"])
def test_device_centre_of_gravity_matches_host(backend, name):
    g, vols = volumes()
    v = vols[name]
    want = oracle.centre_of_gravity(g, v)
    got = backend.centre_of_gravity(backend.upload_image(g, v))
    assert np.max(np.abs(got - want)) <= 1e-4, (got, want)


def test_resident_registration_matches_the_host_path(backend, tmp_path):
    pair = make_pair((48, 52, 44), 2.0, seed=9, disp_max_mm=3.0)
    fixed, moving = [f32(pair.fixed[0])], [f32(pair.moving[0])]
    plan = cohort_plan(1)
    plan.set_linear([30, 15, 8], [4, 2, 1], [2.0, 1.0, 0.0])
    plan.set_syn([10, 6, 4], [4, 2, 1], [2.0, 1.0, 0.0])
    host = register(backend, plan, pair.grid, fixed, pair.grid, moving)
    # the ANTs composite files of the result hold its affine and fields bit for bit
    write_composite(host, pair.grid, tmp_path / "sub_")
    forward = read_composite_h5(tmp_path / "sub_Composite.h5")
    inverse = read_composite_h5(tmp_path / "sub_InverseComposite.h5")
    assert list(forward[0].matrix) == list(host.affine.matrix)
    assert list(forward[0].translation) == list(host.affine.translation)
    assert np.array_equal(forward[1][1], host.forward) and np.array_equal(inverse[0][1], host.inverse)
    assert tuple(forward[1][0].size) == tuple(pair.grid.size) and len(inverse) == 2
    imgs_f = [backend.upload_image(pair.grid, fixed[0])]
    imgs_m = [backend.upload_image(pair.grid, moving[0])]
    dev = register_resident(backend, plan, imgs_f, imgs_m)
    m_h, m_d = np.array(list(host.affine.matrix)), np.array(list(dev.affine.matrix))
    t_h, t_d = np.array(list(host.affine.translation)), np.array(list(dev.affine.translation))
    # the inputs differ by the window's last-ulp roundings and the f32 moment sums (about 1e-5 mm),
    # which the optimiser may amplify through its comparisons: the registration accuracy bounds apply
    assert np.max(np.abs(m_h - m_d)) <= 1e-3, (m_h, m_d)
    assert np.max(np.abs(t_h - t_d)) <= 0.05, (t_h, t_d)
    spacing = 2.0
    for name, want, got in (("forward", host.forward, dev.forward), ("inverse", host.inverse, dev.inverse)):
        d = np.linalg.norm(backend.download_field(got) - want, axis=1) / spacing
        assert d.mean() <= 0.1 and np.percentile(d, 99) <= 0.5, (name, d.mean(), np.percentile(d, 99))
    # the ANTs files of the resident result, written from the device fields, hold their f32 values
    backend.write_composite_fields(tmp_path / "dev_", dev.affine, dev.forward, dev.inverse)
    forward = read_composite_h5(tmp_path / "dev_Composite.h5")
    inverse = read_composite_h5(tmp_path / "dev_InverseComposite.h5")
    assert list(forward[0].matrix) == list(dev.affine.matrix)
    assert np.array_equal(forward[1][1], backend.download_field(dev.forward))
    assert np.array_equal(inverse[0][1], backend.download_field(dev.inverse))
    # the input handles are not consumed: a second run gives the same result
    again = register_resident(backend, plan, imgs_f, imgs_m, fields=False)
    assert np.array_equal(np.array(list(again.affine.matrix)), m_d)
    assert again.forward is None and again.inverse is None
