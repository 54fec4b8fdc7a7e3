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
