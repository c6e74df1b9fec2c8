"""EDI reading and writing (io::read_edi / write_edi / load_edi_directory /
save_edi_directory, src/io/edi_io.cpp:156-619) through the host library.

The reference ships no EDI files, so the reader is pinned by a hand-written
file whose expected values follow from the format itself: impedances in
mV/km/nT times mu0 * 1e3 = 4 pi 1e-4 give ohm, errors are square roots of the
variances times the same factor, DMS coordinates are d + m/60 + s/3600 with
the sign applying to the whole angle, a component without a VAR section gets
the default error (4 pi 1e-4 for Z, 0.1 for the tipper) and a component
without data weight 0.
"""
import math

import numpy as np
import pytest

from maple3dmt_amd import edi
from maple3dmt_amd.native import NativeError

U = 4 * math.pi * 1e-4

SAMPLE = """>HEAD
  DATAID="SITE_07"
  ACQBY="nobody"
  LAT=-35:30:36.0 LONG=+139:15:00.0
  ELEV=212.5

>INFO
  free text the reader skips

>=DEFINEMEAS
  MAXCHAN=7
>EMEAS ID=1001.001 CHTYPE=EX X=0 Y=0

>=MTSECT
  SECTID="SITE_07"
  NFREQ=3

>FREQ ORDER=DEC //3
  1.000000e+02 1.000000e+01
  1.000000e+00
>ZROT //3
  0 0 0
>ZXYR ROT=ZROT //3
  1.0 2.0 3.0
>ZXYI ROT=ZROT //3
  -0.5 -1.0 -1.5
>ZXY.VAR ROT=ZROT //3
  0.04 0.09 0.16
>ZYXR ROT=ZROT //3
  -4.0 -5.0 -6.0
>ZYXI ROT=ZROT //3
  0.25 0.5 0.75
>TXR.EXP //3
  0.1 0.2 0.3
>TXI.EXP //3
  0.01 0.02 0.03
>TX.VAR //3
  0.0004 0.0009 0.0016
>TYR.EXP //3
  -0.1 -0.2 -0.3
>TYI.EXP //3
  0.0 0.0 0.0
>END
"""


def _write(path, text):
    path.write_text(text)
    return str(path)


def test_read_hand_written_sample(tmp_path):
    d = edi.read_edi(_write(tmp_path / "site07.edi", SAMPLE))
    assert d.names == ["SITE_07"]
    lat = -(35 + 30 / 60 + 36 / 3600)
    lon = 139 + 15 / 60
    np.testing.assert_allclose(d.geo[0], [lat, lon, 212.5, 1.0], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(d.freqs, [100.0, 10.0, 1.0])
    zxy = np.array([1.0 - 0.5j, 2.0 - 1.0j, 3.0 - 1.5j]) * U
    zyx = np.array([-4.0 + 0.25j, -5.0 + 0.5j, -6.0 + 0.75j]) * U
    np.testing.assert_allclose(d.z[:, 0, 1], zxy, rtol=1e-15)
    np.testing.assert_allclose(d.z[:, 0, 2], zyx, rtol=1e-15)
    # ZXY has variances, ZYX has none (default error), ZXX/ZYY are absent.
    np.testing.assert_allclose(d.z_err[:, 0, 1], np.array([0.2, 0.3, 0.4]) * U, rtol=1e-15)
    np.testing.assert_allclose(d.z_err[:, 0, 2], U, rtol=1e-15)
    np.testing.assert_array_equal(d.z_w[:, 0], [[0, 1, 1, 0]] * 3)
    np.testing.assert_array_equal(d.z[:, 0, 0], 0)
    # Tipper: .EXP tags, no unit factor, default error 0.1 where VAR is absent.
    np.testing.assert_allclose(d.t[:, 0, 0], [0.1 + 0.01j, 0.2 + 0.02j, 0.3 + 0.03j], rtol=1e-15)
    np.testing.assert_allclose(d.t[:, 0, 1], [-0.1, -0.2, -0.3], rtol=1e-15)
    np.testing.assert_allclose(d.t_err[:, 0, 0], [0.02, 0.03, 0.04], rtol=1e-15)
    np.testing.assert_allclose(d.t_err[:, 0, 1], 0.1, rtol=0)
    np.testing.assert_array_equal(d.t_w[:, 0], [[1, 1]] * 3)


def test_tag_variants(tmp_path):
    """NFREQ only in the MTSECT block, decimal coordinates under LON, TXYR
    tipper names and TXVAR, lowercase tags."""
    text = """>HEAD
DATAID=A1 LAT=12.5 LON=-7.25 ELEV=0
>=MTSECT
NFREQ=2
>FREQ
 5.0 0.5
>zxxr
 1 2
>ZXXI
 3 4
>TXYR.EXP
 0.5 0.6
>TXYI.EXP
 0.0 0.1
>TXVAR.EXP
 0.01 0.04
>END
"""
    d = edi.read_edi(_write(tmp_path / "a1.edi", text))
    assert d.names == ["A1"]
    np.testing.assert_array_equal(d.geo[0], [12.5, -7.25, 0.0, 1.0])
    np.testing.assert_array_equal(d.freqs, [5.0, 0.5])
    np.testing.assert_allclose(d.z[:, 0, 0], np.array([1 + 3j, 2 + 4j]) * U, rtol=1e-15)
    np.testing.assert_array_equal(d.z_w[:, 0], [[1, 0, 0, 0]] * 2)
    np.testing.assert_allclose(d.t[:, 0, 0], [0.5, 0.6 + 0.1j], rtol=1e-15)
    np.testing.assert_allclose(d.t_err[:, 0, 0], [0.1, 0.2], rtol=1e-15)
    np.testing.assert_array_equal(d.t_w[:, 0], [[1, 0]] * 2)


def test_malformed_files_raise(tmp_path):
    with pytest.raises(NativeError, match="cannot open"):
        edi.read_edi(str(tmp_path / "missing.edi"))
    with pytest.raises(NativeError, match="no >FREQ"):
        edi.read_edi(_write(tmp_path / "nofreq.edi", ">HEAD\nDATAID=X\n>END\n"))
    with pytest.raises(NativeError, match="announced"):
        edi.read_edi(_write(tmp_path / "short.edi", ">FREQ //3\n 1 2\n>ZXYR //3\n 1 2 3\n>END\n"))


def _random_data(rng, ns, nf):
    z = (rng.standard_normal((nf, ns, 4)) + 1j * rng.standard_normal((nf, ns, 4))) * 1e-2
    t = (rng.standard_normal((nf, ns, 2)) + 1j * rng.standard_normal((nf, ns, 2))) * 0.2
    z_w = np.ones((nf, ns, 4))
    t_w = np.ones((nf, ns, 2))
    return edi.EDIData(
        names=[f"S{k:02d}" for k in range(ns)],
        geo=np.column_stack([rng.uniform(-60, 60, ns), rng.uniform(-170, 170, ns),
                             rng.uniform(0, 3000, ns), np.ones(ns)]),
        freqs=np.logspace(2, -3, nf), z=z, z_err=np.abs(z) * 0.05 + 1e-6, z_w=z_w,
        t=t, t_err=np.full((nf, ns, 2), 0.03), t_w=t_w)


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    d = _random_data(rng, 1, 11)
    p = str(tmp_path / "rt.edi")
    edi.write_edi(p, d)
    r = edi.read_edi(p)
    assert r.names == d.names
    np.testing.assert_allclose(r.freqs, d.freqs, rtol=1e-6)
    np.testing.assert_allclose(r.geo[0, :3], d.geo[0, :3], rtol=1e-6, atol=1e-6)
    # %.6e text: 7 significant digits on each part.
    np.testing.assert_allclose(r.z.real, d.z.real, rtol=1e-6, atol=0)
    np.testing.assert_allclose(r.z.imag, d.z.imag, rtol=1e-6, atol=0)
    np.testing.assert_allclose(r.z_err, d.z_err, rtol=1e-6)
    np.testing.assert_allclose(r.t.real, d.t.real, rtol=1e-6)
    np.testing.assert_allclose(r.t_err, d.t_err, rtol=1e-6)
    np.testing.assert_array_equal(r.z_w, d.z_w)
    np.testing.assert_array_equal(r.t_w, d.t_w)


def test_directory_union_of_frequencies(tmp_path):
    rng = np.random.default_rng(4)
    d = _random_data(rng, 3, 6)
    edi.save_directory(str(tmp_path), d)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["S00.edi", "S01.edi", "S02.edi"]
    # A fourth station with only two of the frequencies and one new one.
    extra = _random_data(rng, 1, 3)
    extra.names = ["S03"]
    extra.freqs = np.array([d.freqs[1], 7.0, d.freqs[4] * (1 + 1e-9)])
    edi.write_edi(str(tmp_path / "S03.edi"), extra)
    (tmp_path / "notes.txt").write_text("not an EDI file\n")
    r = edi.load_directory(str(tmp_path))
    assert sorted(r.names) == ["S00", "S01", "S02", "S03"]
    want = np.sort(np.concatenate([d.freqs, [7.0]]))[::-1]
    np.testing.assert_allclose(r.freqs, want, rtol=1e-6)
    assert np.all(np.diff(r.freqs) < 0)
    s3 = r.names.index("S03")
    have = {np.argmin(np.abs(r.freqs - f)) for f in extra.freqs}
    for f in range(len(r.freqs)):
        expect = 1.0 if f in have else 0.0
        np.testing.assert_array_equal(r.z_w[f, s3], expect)
    for s, name in enumerate(r.names):
        if name == "S03":
            continue
        k = d.names.index(name)
        rows = [np.argmin(np.abs(r.freqs - f)) for f in d.freqs]
        np.testing.assert_allclose(r.z[rows, s].real, d.z[:, k].real, rtol=1e-6)
        np.testing.assert_array_equal(r.z_w[rows, s], 1.0)
        assert np.all(r.z_w[np.argmin(np.abs(r.freqs - 7.0)), s] == 0.0)
