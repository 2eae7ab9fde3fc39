"""Pins the FP64 quasar-draw oracle and the host set-up of draw_qso (CPU)."""
import numpy as np
import pytest

from oracle import forest_oracle as fo
from oracle import qso_oracle as qo


def test_box_limit_by_brute_force():
    """R_max is the far face's centre distance minus the margin; R_min the
    distance at which the near face's corner cone (with margin) starts; the
    tangents those of the x and y half widths at R_max."""
    LX = LY = 500.0
    LZ, R0, m = 700.0, 3000.0, 6.0
    Rmin, Rmax, tx, ty = qo.box_limit(LX, LY, LZ, R0, m)
    assert Rmax == pytest.approx(R0 + LZ / 2 - m)
    # brute force: the corner direction of half-diagonal (LX/2 - m) at Rmax has
    # cos(alpha) = sqrt(1 - s^2); the near face (z = R0 - LZ/2 + m) meets it at Rmin
    s = (np.hypot(LX, LY) / 2 - m) / Rmax
    zface = R0 - LZ / 2 + m
    assert Rmin == pytest.approx(zface / np.cos(np.arcsin(s)))
    assert tx == pytest.approx(np.tan(np.arcsin((LX / 2 - m) / Rmax)))
    assert ty == pytest.approx(tx)


def test_ra_dec_r_inverts_compute_xyz2_and_matches_the_quadrants():
    rng = np.random.default_rng(0)
    ra0, dec0 = 0.4, -0.3
    ra = rng.uniform(0, 2 * np.pi, 500)
    dec = rng.uniform(-1.2, 1.2, 500)
    R = rng.uniform(100, 5000, 500)
    x, y, z = fo.compute_xyz2(ra, dec, R, ra0, dec0)
    ra2, dec2, R2 = qo.ra_dec_r(x, y, z, ra0, dec0)
    assert np.allclose(ra2, ra, atol=1e-12) and np.allclose(dec2, dec, atol=1e-12) and np.allclose(R2, R)
    # the reference's arctan branches (box.py:215-233) for the four quadrants and the axes
    num = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 1.0, 0.0, -1.0])
    den = np.array([2.0, -2.0, -2.0, 2.0, 1.0, 0.0, -1.0, 0.0])
    want = np.array([np.arctan(0.5), np.arctan(-0.5) + np.pi, np.arctan(0.5) + np.pi, np.arctan(-0.5) + 2 * np.pi,
                     0.0, np.pi / 2, np.pi, 3 * np.pi / 2])
    got = np.arctan2(num, den)
    got = np.where(got < 0, got + 2 * np.pi, got)
    assert np.allclose(got, want, atol=1e-15)


def test_qso_a_of_z_is_a_bias_ratio():
    z = np.linspace(1.5, 3.8, 20)
    for zb in qo.Z_BIAS:
        a = qo.qso_a_of_z(z, zb)
        assert np.allclose(a, (qo.bias_qso(z) / (1 + z)) / (qo.bias_qso(zb) / (1 + zb)))
        assert qo.qso_a_of_z(zb, zb) == pytest.approx(1.0)
        assert np.allclose(a, ((1 + z) / (1 + zb)) ** 0.7)      # b ~ (1+z)^1.7
    assert qo.bias_qso(2.33) == pytest.approx(3.7)


def test_lognormal_mix_returns_a_common_value_and_the_end_fields():
    z = np.linspace(1.0, 4.0, 31)
    for c in (0.0, 0.3, 1.0):
        assert np.allclose(qo.lognormal_mix(z, 2.5, 2.5, 2.5, c), 2.5)
    # at the bias redshifts the interpolations pick one field
    z1, z2, z3 = qo.Z_BIAS
    assert qo.lognormal_mix(z1, 1.0, 2.0, 3.0, 1.0) == pytest.approx(1.0)
    assert qo.lognormal_mix(z2, 1.0, 2.0, 3.0, 0.7) == pytest.approx(2.0)
    assert qo.lognormal_mix(z3, 1.0, 2.0, 3.0, 0.0) == pytest.approx(3.0)


def test_u53_and_philox_words():
    assert qo.u53(0, 0) == 0.0
    assert qo.u53(0xFFFFFFFF, 0xFFFFFFFF) == pytest.approx(1 - 2 ** -53, abs=0)
    assert qo.u53(1 << 31, 0) == 0.5
    u, v = qo.cell_uniforms(7, np.arange(100000), qo.TAG_CELL)
    assert abs(u.mean() - 0.5) < 0.005 and abs(v.mean() - 0.5) < 0.005 and abs(np.corrcoef(u, v)[0, 1]) < 0.01


def _setup(randoms=False, zfix=None):
    from saclaymocks_amd import qso
    nx, ny, nz = 8, 64, 96
    return qso.DrawSetup(nx, ny, nz, 64, 2.19, 2.19, 2.19, 3, 8, sigma=(0.9, 1.0, 1.1), zmin=1.65, zmax=1.78,
                         ra0=0.0, dec0=0.0, dra=5.0, ddec=5.0, zfix=zfix, randoms=randoms)


def test_draw_setup_geometry_and_normalisation():
    from saclaymocks_amd import qso, util
    s = _setup()
    h = util.H
    R0 = h * util.Cosmo().r_comoving(util.Z0)
    assert s.R0 == pytest.approx(R0)
    # slice 3 of 8 of a 64-plane box: x centres of planes 24..31
    assert np.allclose(s.x_axis, (np.arange(8) + 24 + 0.5) * 2.19 - 64 * 2.19 / 2)
    assert np.allclose(s.z_axis, (np.arange(96) + 0.5) * 2.19 + R0 - 96 * 2.19 / 2)
    # the n(z) grid spans the table's bin centres with 2 NZ points, padded by 10 NZ zeros
    t = qso.qso_tables()
    centers = (np.asarray(t["nz_zlow"]) + np.asarray(t["nz_zhigh"])) / 2
    zi = s.dn_z0 + s.dn_dz * np.arange(2 * s.nz)
    assert zi[-1] == pytest.approx(centers[-1])
    assert s.dn_cell.shape == (12 * s.nz,) and np.all(s.dn_cell[2 * s.nz:] == 0)
    assert s.norm > 0 and s.density_max >= s.density_mean > 0
    assert s.z_min == pytest.approx(1.65) and s.z_max == pytest.approx(1.78)


def test_draw_oracle_against_a_cell_loop():
    """The vectorised oracle against a plain loop over the cells of a tiny slab."""
    s = _setup()
    prm = s.params()
    rng = np.random.default_rng(2)
    nx, ny, nz = s.nx, s.ny, s.nz
    p = [rng.normal(0, sg, (nx, ny, nz)).astype(np.float32) for sg in s.sigma]
    vel = [rng.normal(0, 300, (nx, ny, nz)).astype(np.float32) for _ in range(3)]
    # a larger norm so that a handful of cells pass on this small slab
    prm = dict(prm, norm=prm["norm"] * 200, density_max=prm["density_max"] / 50)
    out = qo.draw(p[0], p[1], p[2], vel, prm, seed=11)
    assert out["cond12"] > 0 and len(out["key"]) > 0
    keys = []
    for mz in range(nz):
        for ix in range(nx):
            for iy in range(ny):
                c = (ix * ny + iy) * nz + mz
                r1, r2 = qo.cell_uniforms(11, np.array([c]), qo.TAG_CELL)
                X, Y, Z = prm["x_axis"][ix], prm["y_axis"][iy], prm["z_axis"][mz]
                zc = np.interp(np.sqrt(X * X + Y * Y + Z * Z) / prm["h"], prm["r2z_chi"], prm["r2z_z"])
                coef = np.interp(zc, prm["coef_z"], prm["coef_c"])
                e = [np.float32(np.power(np.exp(np.float32(pk[ix, iy, mz])), np.float32(qo.qso_a_of_z(zc, zb))))
                     for pk, zb in zip(p, qo.Z_BIAS)]
                ptot = qo.lognormal_mix(zc, float(e[0]), float(e[1]), float(e[2]), coef)
                if not r1[0] < prm["norm"] * ptot:
                    continue
                i = int(np.rint((zc - prm["dn_z0"]) / prm["dn_dz"]))
                dens = prm["dn_cell"][i] / qo.variance_correction(zc, prm["sigma"], coef)
                if not prm["density_max"] * r2[0] < dens:
                    continue
                keys.append((mz * nx + ix) * ny + iy)
    # every cond1 & cond2 cell of the loop is a candidate of the oracle; the
    # oracle keeps those that also pass the angle and redshift cut
    assert out["cond12"] == len(keys)
    assert set(out["key"].tolist()) <= set(keys)
    assert np.all(np.diff(out["key"]) > 0)
    assert np.all((out["z_qso_rsd"] > prm["zmin"]) & (out["z_qso_rsd"] < prm["zmax"]))
