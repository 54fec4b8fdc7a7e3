"""Pins of the producer restatement (oracle_tfm_drho and its parts,
TFM.cpp:83-89, zmath.cpp:225-249, SET.cpp:86-99) on the CPU: interpolation
exactness on (bi/tri)linear fields and its clamping, the depth of known ENU
points against an independent FP64 geodetic transform, and the density
perturbation in the cases the model fixes in closed form."""
import numpy as np
import pytest

from ztm_amd import workloads as W

RC = np.array([-15.25, -72.5, 145.0, 300.0, 300.0, 160.0, 10.0, 61.0, 61.0, 33.0])


def test_interp2_reproduces_bilinear_fields_and_clamps(orc):
    di, dj = 7, 5
    j, i = np.meshgrid(np.arange(dj), np.arange(di), indexing="ij")
    f = (1.5 + 0.25 * i - 2.0 * j + 0.125 * i * j).ravel().astype(np.float64)

    def exact(x, y):
        return 1.5 + 0.25 * x - 2.0 * y + 0.125 * x * y

    for x, y in [(0.0, 0.0), (2.5, 1.25), (5.75, 3.5), (6.0, 2.0), (3.0, 4.0), (6.0, 4.0)]:
        assert orc.interp2(f, x, y, di, dj) == pytest.approx(exact(x, y), abs=1e-12)
    # outside the grid the coordinate is clamped to [0, d-1]
    assert orc.interp2(f, -3.0, 1.0, di, dj) == pytest.approx(exact(0.0, 1.0), abs=1e-12)
    assert orc.interp2(f, 9.0, 7.0, di, dj) == pytest.approx(exact(6.0, 4.0), abs=1e-12)


def test_interp3_reproduces_trilinear_fields_and_clamps_at_d_minus_2(orc):
    di, dj, dk = 5, 4, 6
    k, j, i = np.meshgrid(np.arange(dk), np.arange(dj), np.arange(di), indexing="ij")
    f = (2.0 + i - 0.5 * j + 0.25 * k + 0.1 * i * j * k).ravel().astype(np.float64)

    def exact(x, y, z):
        return 2.0 + x - 0.5 * y + 0.25 * z + 0.1 * x * y * z

    for p in [(0.0, 0.0, 0.0), (1.5, 1.25, 3.75), (3.0, 2.0, 4.0), (2.9, 1.1, 0.5)]:
        assert orc.interp3(f, *p, di, dj, dk) == pytest.approx(exact(*p), abs=1e-12)
    # the reference clamps to [0, d-2] (zmath.cpp:239-241): beyond d-2 the
    # value is that of the d-2 plane, the last plane is never weighted
    assert orc.interp3(f, 3.5, 1.0, 1.0, di, dj, dk) == pytest.approx(exact(3.0, 1.0, 1.0))
    assert orc.interp3(f, 7.0, 1.0, 5.0, di, dj, dk) == pytest.approx(exact(3.0, 1.0, 4.0))


def test_c2g_depth_against_fp64_geodetic_transform(orc):
    # the ENU origin of the grid is the reference point at depth RC[2]
    assert orc.c2g_depth(RC, 30.0, 30.0, 16.0) == pytest.approx(145.0, abs=1e-3)
    rng = np.random.default_rng(3)
    lat = RC[0] + rng.uniform(-2.5, 2.5, 40)
    lon = RC[1] + rng.uniform(-2.5, 2.5, 40)
    dep = rng.uniform(-5.0, 300.0, 40)
    x, y, z = W._geo2enu(RC[:3], lat, lon, dep)
    for a in range(40):
        c = ((x[a] + RC[3]) / RC[6], (y[a] + RC[4]) / RC[6], (z[a] + RC[5]) / RC[6])
        # float sqrt of |ECEF|^2 ~ 4e7 km^2: about 2e-4 km of rounding
        assert orc.c2g_depth(RC, *c) == pytest.approx(dep[a], abs=1e-3), a
    # straight up from the origin: depth falls by one cell edge per k step
    d0 = orc.c2g_depth(RC, 30.0, 30.0, 16.0)
    assert orc.c2g_depth(RC, 30.0, 30.0, 17.0) == pytest.approx(d0 - 10.0, abs=1e-3)


def _flat(di, dj, dk, sea_k, moho_k, v):
    sea = np.full(di * dj, sea_k)
    moho = np.full(di * dj, moho_k)
    vel = np.full(di * dj * dk, v)
    return sea, moho, vel


def test_drho_closed_form_cases(orc):
    """Flat surfaces on a small grid near the reference point. Reference Moho
    m0 = sea - 3.5 = 6.5, rendered Moho m1 = 4.0, band half-width li = 1/fv."""
    di, dj, dk, fh, fv = 6, 5, 12, 1, 2
    rc = np.array([-15.25, -72.5, 0.0, 25.0, 20.0, 110.0, 10.0])   # top k=11 at +0 km
    vm, ro, dr, v = 8.0, -0.4, 0.3, 8.2
    sea, moho, vel = _flat(di, dj, dk, 10.0, 4.0, v)
    drho, nn = orc.tfm_drho(di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr)
    ii, jj = (di - 1) * fh + 1, (dj - 1) * fh + 1
    k = (np.arange(drho.size) // (ii * jj)) / fv
    depth = np.array([orc.c2g_depth(rc, (n % ii) / fh, ((n // ii) % jj) / fh, k[n])
                      for n in range(drho.size)])
    li = 1.0 / fv
    below = (k < 4.0 - li) & (depth > 10)                  # under both Mohos
    root = (k > 4.0 + li) & (k < 6.5 - li) & (depth > 10)  # crust where the reference has mantle
    crust = k > 6.5 + li                                   # crust in both: r0 = r1 = ro
    assert below.any() and root.any() and crust.any()
    np.testing.assert_allclose(drho[below], dr * (v - vm), rtol=1e-14)
    np.testing.assert_allclose(drho[root], ro, rtol=1e-14)
    assert np.all(drho[crust] == 0)
    # reference band centre k = m0 = 6.5: r0 = ro / 2, r1 = ro
    at_m0 = (k == 6.5) & (depth > 10)
    np.testing.assert_allclose(drho[at_m0], ro - ro / 2, rtol=1e-14)
    # rendered band: r1 + (k - m1 + li) (r1 + ro) / (2 li) as TFM.cpp:88 writes
    # it; at k = 4.0 (t = 1/2) with r1 = dr (v - vm) = 0.06
    r1 = dr * (v - vm)
    at_m1 = (k == 4.0) & (depth > 10)
    np.testing.assert_allclose(drho[at_m1], r1 + 0.5 * (r1 + ro), rtol=1e-14)
    # the kept list is the non-zero cells, ascending, all deeper than 10 km
    assert np.array_equal(nn, np.flatnonzero(drho))
    assert np.all(depth[nn] > 10)
    # shallow cells are dropped even where r0 != r1
    shallow = (depth <= 10) & (k < 4.0 - li)
    assert np.all(drho[shallow] == 0)


def test_drho_threshold_and_reference_workload(orc):
    w = W.tfm_fm()
    drho, nn = orc.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.vm,
                            w.ro, w.dr)
    assert drho.size == w.n_fine == 61 * 61 * 65
    kept = drho != 0
    assert np.all(np.abs(drho[kept]) > 0.001) and np.array_equal(nn, np.flatnonzero(kept))
    # a crustal root (negative) under the topography and a mantle signal
    assert drho.min() < -0.3 and drho.max() > 0.0
    # v = vm everywhere leaves only the Moho geometry: |drho| <= |ro|; ro (a
    # crustal root) between a deeper rendered Moho and the reference one,
    # positive values only in columns whose rendered Moho is shallower
    drho0, _ = orc.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho,
                            np.full_like(w.vel, w.vm), w.vm, w.ro, w.dr)
    assert np.all(np.abs(drho0) <= abs(w.ro) + 1e-12)
    assert np.isin(w.ro, np.round(drho0[drho0 != 0], 12))
    ii = (w.di - 1) * w.fh + 1
    jj = (w.dj - 1) * w.fh + 1
    n = np.flatnonzero(drho0 > 0)
    col = np.rint((n % ii) / w.fh).astype(int) + w.di * np.rint(((n // ii) % jj) / w.fh).astype(int)
    shallower = w.moho > w.sea - 3.5
    assert n.size > 0 and np.all(shallower[col] | np.roll(shallower, 1)[col]
                                 | np.roll(shallower, -1)[col]
                                 | np.roll(shallower, w.di)[col] | np.roll(shallower, -w.di)[col])
