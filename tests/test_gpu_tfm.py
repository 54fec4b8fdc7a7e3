"""The gravity FM of one ZTM forward-model call on the device (TFM.cpp:83-94):
the rendered density perturbation equals the FP64 restatement's value
narrowed to float bit for bit, and the in-process gravity equals the FP64
oracle of GF on that field (1e-6 of max |g|, as the reference-ABI tests) and
the reference's CPU path on the double field (1e-5)."""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def device_drho(w, dev):
    import torch
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    return gfm.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, t(w.sea), t(w.moho), t(w.vel), w.vm,
                        w.ro, w.dr).cpu().numpy()


@pytest.mark.parametrize("geom", [dict(), dict(fh=2, fv=1, di=23, dj=19, dk=14),
                                  dict(fh=1, fv=3, di=31, dj=27, dk=17)])
def test_drho_bitwise_equals_restatement(dev, orc, geom):
    w = W.tfm_fm(**geom)
    want, nn = orc.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.vm,
                            w.ro, w.dr)
    got = device_drho(w, dev)
    assert got.dtype == np.float32 and got.size == want.size
    assert np.array_equal(got, want.astype(np.float32))
    assert nn.size > 1000


def test_gravity_fm_host_matches_oracles(dev, orc):
    w = W.tfm_fm()
    G, drho = gfm.tfm_gravity_host(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho,
                                   w.vel, w.surf, w.vm, w.ro, w.dr, want_drho=True)
    want_d, nn = orc.tfm_drho(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.vm,
                              w.ro, w.dr)
    assert np.array_equal(drho, want_d.astype(np.float32))
    # GF on the float field (the reference's GPU branch, TFM.cpp:90-94)
    g_ref = orc.gf(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, drho, w.surf.astype(np.float32))
    assert relerr(G, g_ref) < 1e-6
    # the CPU branch (TFM.cpp:95-99) on the double field and its kept list
    g_cpu = orc.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, want_d, w.surf, nn, threads=16)
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    assert relerr(G[interior], g_cpu[interior]) < 1e-5
    assert np.all(G[~interior] == 0)
    assert np.abs(G).max() > 10.0     # mGal-scale signal of a crustal root


def test_gravity_fm_repeated_calls_and_errors(dev, orc):
    a = W.tfm_fm(seed=W.SEED0 + 16)
    b = W.tfm_fm(seed=W.SEED0 + 26)
    args = lambda w: (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel, w.surf,
                      w.vm, w.ro, w.dr)
    Ga1 = gfm.tfm_gravity_host(*args(a))
    Gb = gfm.tfm_gravity_host(*args(b))
    Ga2 = gfm.tfm_gravity_host(*args(a))
    assert np.array_equal(Ga1, Ga2) and not np.array_equal(Ga1, Gb)
    w = a
    with pytest.raises(gfm.ZtmError):
        gfm.tfm_gravity_host(w.bfr, 1, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea[:w.dj],
                             w.moho[:w.dj], w.vel[:w.dj * w.dk], w.surf[:w.dj], w.vm, w.ro, w.dr)
    bad_rc = w.rc.copy()
    bad_rc[6] = np.nan
    with pytest.raises(gfm.ZtmError):
        gfm.tfm_gravity_host(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, bad_rc, w.sea, w.moho, w.vel,
                             w.surf, w.vm, w.ro, w.dr)


def test_batched_variants_match_single_calls(dev, orc):
    """Six FM variants (gravity parameters and Moho perturbed per variant) in
    one call: drho bitwise equal to the single-call producer, gravity within
    5e-6 of the single FP32 path and of the FP64 oracle (split-BF16 MFMA)."""
    import torch
    w = W.tfm_fm()
    B, M = 6, w.di * w.dj
    params = np.array([[w.vm + 0.01 * b, w.ro - 0.005 * b, w.dr + 0.002 * b] for b in range(B)])
    bump = np.sin(np.arange(M) * 0.37)
    moho = np.stack([w.moho + 0.1 * b * bump for b in range(B)])
    t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    g, drho = gfm.tfm_gravity_batched(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, t64(w.sea),
                                      t64(moho), t64(w.vel), t64(params), s32)
    g, drho = g.cpu().numpy(), drho.cpu().numpy()
    for b in range(B):
        Gb, db = gfm.tfm_gravity_host(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, moho[b],
                                      w.vel, w.surf, *params[b], want_drho=True)
        assert np.array_equal(drho[b], db), b
        assert relerr(g[b], Gb) < 5e-6, b
        assert np.all(g[b][Gb == 0] == 0)
    want = orc.gf(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, drho[B - 1], w.surf.astype(np.float32))
    assert relerr(g[B - 1], want) < 5e-6
    # shared Moho and velocity (stride 0): three gravity-parameter columns
    g3, _ = gfm.tfm_gravity_batched(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, t64(w.sea),
                                    t64(w.moho), t64(w.vel), t64(params[:3]), s32)
    g3 = g3.cpu().numpy()
    G0 = gfm.tfm_gravity_host(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel,
                              w.surf, *params[2])
    assert relerr(g3[2], G0) < 5e-6


def test_batched_field_counts_match_single_calls(dev, orc):
    """5 to 16 FM variants per call, which run the multi-field forward with
    5-8 fields per pass (two-row groups) and two passes past 8: every
    variant's gravity equals its own one-variant call (one field per pass,
    four-row groups) to 1e-12, since each row's FP32 sum is formed the same
    way and only the FP64 grouping of rows into slices differs, and the
    last variant of the largest call is within 5e-6 of the FP64 oracle."""
    import torch
    w = W.tfm_fm()
    M = w.di * w.dj
    Bmax = 16
    params = np.array([[w.vm + 0.01 * b, w.ro - 0.004 * b, w.dr + 0.003 * b]
                       for b in range(Bmax)])
    bump = np.cos(np.arange(M) * 0.21)
    moho = np.stack([w.moho + 0.05 * b * bump for b in range(Bmax)])
    t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    args = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, t64(w.sea))
    single = [gfm.tfm_gravity_batched(*args, t64(moho[b:b + 1]), t64(w.vel),
                                      t64(params[b:b + 1]), s32)[0].cpu().numpy()[0]
              for b in range(Bmax)]
    for B in (5, 6, 7, 8, 12, 16):
        print(f"[tfm] {B} variants in one call", flush=True)
        g, _ = gfm.tfm_gravity_batched(*args, t64(moho[:B]), t64(w.vel), t64(params[:B]), s32)
        g = g.cpu().numpy()
        for b in range(B):
            assert relerr(g[b], single[b]) < 1e-12, (B, b)
    want = gfm.tfm_gravity_host(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, moho[-1],
                                w.vel, w.surf, *params[-1], want_drho=True)[1]
    want = orc.gf(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, want, w.surf.astype(np.float32))
    assert relerr(g[-1], want) < 5e-6


def _t64(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)


@pytest.mark.parametrize("geom", [dict(), dict(fh=2, fv=1, di=23, dj=19, dk=14)])
def test_drho_jvp_matches_central_differences(dev, orc, geom):
    """d drho / d(vm, ro, dr) against central differences of the FP64
    restatement of TFM.cpp:86-89 on the cells that stay in one linear piece
    and stay kept over [p - h, p + h] (there the difference is exact up to
    rounding); cells the producer or GF drops at p must give exactly 0."""
    w = W.tfm_fm(**geom)
    args = (w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea, w.moho, w.vel)
    dd = gfm.tfm_drho_jvp(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev),
                          _t64(w.moho, dev), _t64(w.vel, dev), w.vm, w.ro, w.dr).cpu().numpy()
    p0 = np.array([w.vm, w.ro, w.dr])
    d0, _ = orc.tfm_drho(*args, *p0)
    kept = np.abs(d0.astype(np.float32)) > np.float32(0.001)
    h = 1e-4
    for q in range(3):
        e = np.zeros(3)
        e[q] = h
        dp, _ = orc.tfm_drho(*args, *(p0 + e))
        dm, _ = orc.tfm_drho(*args, *(p0 - e))
        right, left = (dp - d0) / h, (d0 - dm) / h
        keep3 = kept & (np.abs(dp.astype(np.float32)) > 0.001) & (np.abs(dm.astype(np.float32)) > 0.001)
        smooth = keep3 & (np.abs(right - left) <= 1e-7 * (1 + np.abs(right)))
        fd = (dp - dm) / (2 * h)
        assert smooth.sum() > 1000, q
        scale = max(np.abs(fd[smooth]).max(), 1e-30)
        assert np.max(np.abs(dd[q][smooth] - fd[smooth])) / scale < 1e-6, q
        assert np.all(dd[q][~kept] == 0), q
    # somewhere the ro derivative is the band fraction, not just 0 or +-1
    band = kept & (np.abs(dd[1]) > 1e-3) & (np.abs(np.abs(dd[1]) - 1) > 1e-3)
    assert band.sum() > 0


def test_gravity_jvp_matches_gf_oracle(dev, orc):
    """dG/d(vm, ro, dr): the FP32 GF forward of each derivative field
    against the FP64 GF oracle on the same field (scaled by 2^20, exact in
    float, so that GF's .001 skip drops none of them), 1e-6 as the
    reference-ABI tests; exterior stations 0."""
    import torch
    w = W.tfm_fm()
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    dg, dd = gfm.tfm_gravity_jvp(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev),
                                 _t64(w.moho, dev), _t64(w.vel, dev), w.vm, w.ro, w.dr, s32)
    dg, dd = dg.cpu().numpy(), dd.cpu().numpy()
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    c = np.float32(2.0 ** 20)
    for q in range(3):
        want = orc.gf(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, dd[q] * c, w.surf.astype(np.float32))
        want = np.asarray(want, np.float64) / float(c)
        assert np.abs(want).max() > 0, q
        assert relerr(dg[q], want) < 1e-6, q
        assert np.all(dg[q][~interior] == 0), q
    # the fields equal the stand-alone derivative call bit for bit
    dd2 = gfm.tfm_drho_jvp(w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev),
                           _t64(w.moho, dev), _t64(w.vel, dev), w.vm, w.ro, w.dr).cpu().numpy()
    assert np.array_equal(dd, dd2)


def test_gravity_vjp_is_the_adjoint_of_the_jvp(dev, orc):
    """r . (dG/db) from the adjoint path (G^T r dotted with the derivative
    fields) equals the same sum formed from the forward columns, for a
    seeded residual (exterior entries must not matter)."""
    import torch
    w = W.tfm_fm()
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    args = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev), _t64(w.moho, dev),
            _t64(w.vel, dev), w.vm, w.ro, w.dr, s32)
    dg, dd = gfm.tfm_gravity_jvp(*args)
    r = W.normal(W.SEED0 + 31, 1, w.di * w.dj, 1.0)
    out, dd2 = gfm.tfm_gravity_vjp(*args, _t64(r, dev))
    out, dg = out.cpu().numpy(), dg.cpu().numpy()
    assert np.array_equal(dd.cpu().numpy(), dd2.cpu().numpy())
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    for q in range(3):
        terms = r[interior] * dg[q][interior]
        assert abs(out[q] - terms.sum()) <= 1e-6 * np.abs(terms).sum(), q
    # exterior residuals are ignored
    r2 = r.copy()
    r2[~interior] = 1e3
    out2, _ = gfm.tfm_gravity_vjp(*args, _t64(r2, dev))
    assert np.array_equal(out2.cpu().numpy(), out)


def _tangents(w, nd, seed=61):
    """nd smooth tangent directions of the Moho (k units) and velocity
    (km/s) grids: direction 0 moves only the Moho, 1 only the velocity, the
    rest both."""
    k, j, i = np.meshgrid(np.arange(w.dk), np.arange(w.dj), np.arange(w.di), indexing="ij")
    ph = W.uniform(seed, 1, 4 * nd, 0.0, 2 * np.pi).reshape(nd, 4)
    dmo = np.stack([0.3 * np.sin((0.09 + 0.01 * (b % 5)) * i[0] + 0.07 * j[0] + ph[b, 0])
                    .ravel() for b in range(nd)])
    dve = np.stack([0.05 * np.cos(0.08 * i + (0.05 + 0.01 * (b % 3)) * j + 0.2 * k + ph[b, 1])
                    .ravel() for b in range(nd)])
    dmo[1] = 0.0
    dve[0] = 0.0
    return dmo, dve


def _fd_drho(orc, w, dmo, dve, h):
    base = (w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea)
    p = (w.vm, w.ro, w.dr)
    d0, _ = orc.tfm_drho(*base, w.moho, w.vel, *p)
    dp, _ = orc.tfm_drho(*base, w.moho + h * dmo, w.vel + h * dve, *p)
    dm, _ = orc.tfm_drho(*base, w.moho - h * dmo, w.vel - h * dve, *p)
    return d0, dp, dm


@pytest.mark.parametrize("geom", [dict(), dict(fh=2, fv=1, di=23, dj=19, dk=14)])
def test_grid_jvp_fields_match_central_differences(dev, orc, geom):
    """d drho along tangent Moho / velocity grids against central differences
    of the FP64 restatement of TFM.cpp:83-89, on the cells that stay in one
    linear piece and stay kept over [-h, h] (there the difference is exact up
    to rounding): 1e-6 of the largest; dropped cells exactly 0; the Moho
    direction is non-zero only in the band, the velocity one also below it."""
    w = W.tfm_fm(**geom)
    dmo, dve = _tangents(w, 3)
    dg, dd = gfm.tfm_gravity_jvp_grids(
        w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev), _t64(w.moho, dev),
        _t64(w.vel, dev), w.vm, w.ro, w.dr,
        __import__("torch").from_numpy(w.surf.astype(np.float32)).to(dev),
        dmoho=_t64(dmo, dev), dvel=_t64(dve, dev))
    dd = dd.cpu().numpy()
    h = 1e-4
    for b in range(3):
        d0, dp, dm = _fd_drho(orc, w, dmo[b], dve[b], h)
        kept = np.abs(d0.astype(np.float32)) > np.float32(0.001)
        right, left = (dp - d0) / h, (d0 - dm) / h
        keep3 = kept & (np.abs(dp.astype(np.float32)) > 0.001) & \
            (np.abs(dm.astype(np.float32)) > 0.001)
        smooth = keep3 & (np.abs(right - left) <= 1e-7 * (1 + np.abs(right)))
        fd = (dp - dm) / (2 * h)
        assert smooth.sum() > 500, b
        scale = max(np.abs(fd[smooth]).max(), 1e-30)
        assert scale > 1e-3, b
        assert np.max(np.abs(dd[b][smooth] - fd[smooth])) / scale < 1e-6, b
        assert np.all(dd[b][~kept] == 0), b
    nz0 = np.count_nonzero(dd[0])
    assert 0 < nz0 < np.count_nonzero(dd[1]), "Moho tangents act only in the band"


def test_grid_jvp_gravity_matches_gf_oracle(dev, orc):
    """dG of tangent fields: 3 directions (one FP32 GF forward with FP64
    near-field correction, 1e-6) and 40 (split MFMA ensemble, 5e-6) against
    the FP64 GF oracle on the same fields (scaled by 2^20, exact in float, so
    GF's .001 skip drops none); exterior stations 0; the shared directions'
    fields are equal bit for bit between the two calls."""
    import torch
    w = W.tfm_fm()
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    args = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev), _t64(w.moho, dev),
            _t64(w.vel, dev), w.vm, w.ro, w.dr, s32)
    dmo, dve = _tangents(w, 40)
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    c = np.float32(2.0 ** 20)
    fields = {}
    for nd, tol, check in ((3, 1e-6, range(3)), (40, 5e-6, (0, 1, 2, 17, 39))):
        dg, dd = gfm.tfm_gravity_jvp_grids(*args, dmoho=_t64(dmo[:nd], dev),
                                           dvel=_t64(dve[:nd], dev))
        dg, dd = dg.cpu().numpy(), dd.cpu().numpy()
        fields[nd] = dd
        for b in check:
            want = orc.gf(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, dd[b] * c,
                          w.surf.astype(np.float32))
            want = np.asarray(want, np.float64) / float(c)
            assert np.abs(want).max() > 0, (nd, b)
            assert relerr(dg[b], want) < tol, (nd, b)
            assert np.all(dg[b][~interior] == 0), (nd, b)
    assert np.array_equal(fields[3], fields[40][:3])


def test_grid_jvp_matches_gravity_central_differences(dev, orc):
    """End to end: dG along one tangent of both grids against central
    differences (h = 1e-6) of the whole FP64 CPU gravity path (producer
    TFM.cpp:83-89 and gravity TFM.cpp:95-99) where no kept cell changes piece
    or filter within the step (checked), 1e-5 of the largest."""
    import torch
    w = W.tfm_fm()
    dmo, dve = _tangents(w, 3)
    dmo, dve = dmo[2:3], dve[2:3]
    h = 1e-6
    d0, dp, dm = _fd_drho(orc, w, dmo[0], dve[0], h)
    kept = np.abs(d0.astype(np.float32)) > np.float32(0.001)
    keep3 = kept & (np.abs(dp.astype(np.float32)) > 0.001) & (np.abs(dm.astype(np.float32)) > 0.001)
    right, left = (dp - d0) / h, (d0 - dm) / h
    assert np.array_equal(keep3, kept)
    assert np.all(np.abs(right - left)[kept] <= 1e-6 * (1 + np.abs(right[kept])))
    base = (w.di, w.dj, w.dk, w.fh, w.fv, w.rc, w.sea)
    _, nnp = orc.tfm_drho(*base, w.moho + h * dmo[0], w.vel + h * dve[0], w.vm, w.ro, w.dr)
    _, nnm = orc.tfm_drho(*base, w.moho - h * dmo[0], w.vel - h * dve[0], w.vm, w.ro, w.dr)
    assert np.array_equal(nnp, nnm)
    surf = w.surf
    Gp = orc.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, dp, surf, nnp)
    Gm = orc.tfm_gravity(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, dm, surf, nnm)
    fd = (Gp - Gm) / (2 * h)
    s32 = torch.from_numpy(surf.astype(np.float32)).to(dev)
    dg, _ = gfm.tfm_gravity_jvp_grids(w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc,
                                      _t64(w.sea, dev), _t64(w.moho, dev), _t64(w.vel, dev),
                                      w.vm, w.ro, w.dr, s32, dmoho=_t64(dmo, dev),
                                      dvel=_t64(dve, dev))
    assert np.abs(fd).max() > 0
    assert relerr(dg.cpu().numpy()[0], fd) < 1e-5


def test_grid_vjp_is_the_adjoint_of_the_jvp(dev, orc):
    """d(r . G)/dM and d(r . G)/dV from one adjoint: <J^T r, v> = <r, J v>
    for three tangent directions (1e-6 of the sum of |terms|); node by node,
    the gradient of 8 Moho nodes and 8 velocity nodes equals r . dG along
    that node's unit direction (one 16-direction JVP call, 1e-5); exterior
    residuals are ignored."""
    import torch
    w = W.tfm_fm()
    M = w.di * w.dj
    s32 = torch.from_numpy(w.surf.astype(np.float32)).to(dev)
    args = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc, _t64(w.sea, dev), _t64(w.moho, dev),
            _t64(w.vel, dev), w.vm, w.ro, w.dr, s32)
    r = W.normal(W.SEED0 + 33, 1, M, 1.0)
    gm, gv = gfm.tfm_gravity_vjp_grids(*args, _t64(r, dev))
    gm, gv = gm.cpu().numpy(), gv.cpu().numpy()
    interior = np.zeros((w.dj, w.di), bool)
    interior[w.bfr:w.dj - w.bfr, w.bfr:w.di - w.bfr] = True
    interior = interior.ravel()
    dmo, dve = _tangents(w, 3)
    dg, _ = gfm.tfm_gravity_jvp_grids(*args, dmoho=_t64(dmo, dev), dvel=_t64(dve, dev))
    dg = dg.cpu().numpy()
    for b in range(3):
        terms = r[interior] * dg[b][interior]
        lhs = gm @ dmo[b] + gv @ dve[b]
        assert abs(lhs - terms.sum()) <= 1e-6 * np.abs(terms).sum(), b
    mn = np.argsort(-np.abs(gm))[:4].tolist() + \
        np.flatnonzero(gm)[(W.uniform(33, 2, 4) * np.count_nonzero(gm)).astype(int)].tolist()
    vn = np.argsort(-np.abs(gv))[:4].tolist() + \
        np.flatnonzero(gv)[(W.uniform(33, 3, 4) * np.count_nonzero(gv)).astype(int)].tolist()
    em = np.zeros((16, M))
    ev = np.zeros((16, M * w.dk))
    for q, node in enumerate(mn):
        em[q, node] = 1.0
    for q, node in enumerate(vn):
        ev[8 + q, node] = 1.0
    dg1, _ = gfm.tfm_gravity_jvp_grids(*args, dmoho=_t64(em, dev), dvel=_t64(ev, dev))
    dg1 = dg1.cpu().numpy()
    for q in range(16):
        terms = r[interior] * dg1[q][interior]
        got = gm[mn[q]] if q < 8 else gv[vn[q - 8]]
        assert abs(got - terms.sum()) <= 1e-5 * max(np.abs(terms).sum(), 1e-30), q
    r2 = r.copy()
    r2[~interior] = 1e3
    gm2, gv2 = gfm.tfm_gravity_vjp_grids(*args, _t64(r2, dev))
    assert np.array_equal(gm2.cpu().numpy(), gm) and np.array_equal(gv2.cpu().numpy(), gv)
