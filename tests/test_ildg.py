"""ILDG/LIME reader and writer (libgfhip host code) against an independent numpy decoder.

Fixture: tests/golden/conf_08080808.ildg, the SU(3) 8^4 configuration the reference ships
with its measurement samples (samples/measurements/conf_08080808.ildg).  It holds a single
ildg-binary-data record (no ildg-format record), so the reader must infer 64-bit precision
from the payload size.  The reference records no plaquette for it; its value is pinned here
by the numpy brute force and by physics: a thermalized SU(3) Wilson configuration at
beta = 5.7 has an average plaquette near 0.55, while decoding with the matrix transposed or
the site order reversed gives values near 0.06 and 0.
"""
import os

import numpy as np
import pytest

import lattice_tools as lt

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "conf_08080808.ildg")
DIMS = (8, 8, 8, 8)


# ---------------------------------------------------------------- independent decoder


def lime_records(raw):
    """[(type, mb, me, payload bytes)] of a LIME byte string."""
    out, pos = [], 0
    while pos < len(raw):
        h = raw[pos:pos + 144]
        assert int.from_bytes(h[0:4], "big") == 0x456789AB
        flags = int.from_bytes(h[6:8], "big")
        n = int.from_bytes(h[8:16], "big")
        typ = h[16:144].split(b"\0", 1)[0].decode()
        out.append((typ, bool(flags & 0x8000), bool(flags & 0x4000), raw[pos + 144:pos + 144 + n]))
        pos += 144 + (n + 7) // 8 * 8
    return out


def decode_payload(payload, dims, width):
    """payload in file order (t, z, y, x, mu, row, col) -> (4, V, 9) with element i + 3 j"""
    nx, ny, nz, nt = dims
    d = np.frombuffer(payload, dtype=">f8" if width == 8 else ">f4").astype(np.float64)
    M = (d[0::2] + 1j * d[1::2]).reshape(nt, nz, ny, nx, 4, 3, 3)
    M = np.transpose(M, (4, 0, 1, 2, 3, 5, 6)).reshape(4, -1, 3, 3)  # site x fastest
    return np.ascontiguousarray(np.swapaxes(M, -1, -2)).reshape(4, -1, 9)


@pytest.fixture(scope="module")
def ildg(gf):
    from gaugefields_jl_amd import ildg as m

    return m


@pytest.fixture(scope="module")
def fixture_raw():
    with open(FIXTURE, "rb") as f:
        return f.read()


def test_fixture_read_matches_numpy(ildg, fixture_raw):
    recs = lime_records(fixture_raw)
    assert [r[0] for r in recs] == ["ildg-binary-data"]
    expect = decode_payload(recs[0][3], DIMS, 8)
    U, prec = ildg.read_ildg(FIXTURE, DIMS)
    assert prec == 64
    assert np.array_equal(U, expect)
    info = ildg.ildg_info(FIXTURE)
    assert info == {"L": None, "precision": None, "payload_bytes": 8 ** 4 * 4 * 18 * 8}


def test_fixture_is_thermalized_su3(ildg, oracle):
    U, _ = ildg.read_ildg(FIXTURE, DIMS)
    m = U.reshape(-1, 3, 3, order="C").transpose(0, 2, 1)  # element (i, j) at i + 3 j
    assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(m) - 1).max() < 1e-13
    V = 8 ** 4
    p = oracle.plaquette_sum(DIMS, U) / (6 * V * 3)
    assert p == pytest.approx(lt.plaquette_field_sum(U, DIMS) / (6 * V * 3), rel=1e-13)
    assert 0.54 < p < 0.57
    # the window rejects the plausible decoding mistakes
    Ut = np.ascontiguousarray(U.reshape(4, -1, 3, 3).swapaxes(-1, -2)).reshape(4, -1, 9)
    assert not 0.54 < oracle.plaquette_sum(DIMS, Ut) / (6 * V * 3) < 0.57
    Ur = np.ascontiguousarray(U.reshape(4, 8, 8, 8, 8, 9).transpose(0, 4, 3, 2, 1, 5)).reshape(4, -1, 9)
    assert not 0.54 < oracle.plaquette_sum(DIMS, Ur) / (6 * V * 3) < 0.57


def test_roundtrip_64_reproduces_payload(ildg, fixture_raw, tmp_path):
    U, _ = ildg.read_ildg(FIXTURE, DIMS)
    out = tmp_path / "rt64.ildg"
    ildg.write_ildg(out, DIMS, U, 64)
    recs = lime_records(out.read_bytes())
    assert [(r[0], r[1], r[2]) for r in recs] == [("ildg-format", True, False), ("ildg-binary-data", False, True)]
    xml = recs[0][3].decode()
    for tag, v in (("precision", 64), ("lx", 8), ("ly", 8), ("lz", 8), ("lt", 8), ("field", "su3gauge")):
        assert f"<{tag}>{v}</{tag}>" in xml
    assert recs[1][3] == lime_records(fixture_raw)[0][3]
    U2, prec = ildg.read_ildg(out, DIMS)
    assert prec == 64 and np.array_equal(U2, U)
    assert ildg.ildg_info(out) == {"L": DIMS, "precision": 64, "payload_bytes": len(recs[1][3])}


def test_roundtrip_32_rounds_to_single(ildg, tmp_path):
    # test/ildg_io.jl:57: the 32-bit file restores ComplexF64.(ComplexF32.(original))
    U, _ = ildg.read_ildg(FIXTURE, DIMS)
    out = tmp_path / "rt32.ildg"
    ildg.write_ildg(out, DIMS, U, 32)
    recs = lime_records(out.read_bytes())
    assert len(recs[1][3]) == 8 ** 4 * 4 * 18 * 4
    U2, prec = ildg.read_ildg(out, DIMS)
    assert prec == 32
    assert np.array_equal(U2, U.astype(np.complex64).astype(np.complex128))
    assert np.array_equal(decode_payload(recs[1][3], DIMS, 4), U2)


@pytest.mark.parametrize("dims", [(2, 3, 4, 5), (1, 1, 1, 1)])
def test_roundtrip_ragged_extents(ildg, tmp_path, dims):
    rng = np.random.default_rng(11)
    V = int(np.prod(dims))
    U = (rng.normal(size=(4, V, 9)) + 1j * rng.normal(size=(4, V, 9)))
    out = tmp_path / "r.ildg"
    ildg.write_ildg(out, dims, U, 64)
    U2, _ = ildg.read_ildg(out, dims)
    assert np.array_equal(U2, U)
    assert np.array_equal(decode_payload(lime_records(out.read_bytes())[1][3], dims, 8), U)


def test_errors(ildg, gf, tmp_path):
    GFError = gf.GFError
    U = np.zeros((4, 16, 9), np.complex128)
    with pytest.raises(GFError, match="precision must be 32 or 64"):
        ildg.write_ildg(tmp_path / "p16.ildg", (2, 2, 2, 2), U, 16)
    with pytest.raises(GFError, match="cannot open"):
        ildg.read_ildg(tmp_path / "missing.ildg", (2, 2, 2, 2))
    with pytest.raises(GFError, match="does not match"):
        ildg.read_ildg(FIXTURE, (8, 8, 8, 4))
    good = tmp_path / "g.ildg"
    ildg.write_ildg(good, (2, 2, 2, 2), U, 64)
    with pytest.raises(GFError, match="extents differ"):
        ildg.read_ildg(good, (2, 2, 4, 1))
    raw = good.read_bytes()
    (tmp_path / "trunc.ildg").write_bytes(raw[:-100])
    with pytest.raises(GFError, match="does not match|truncated"):
        ildg.read_ildg(tmp_path / "trunc.ildg", (2, 2, 2, 2))
    (tmp_path / "junk.ildg").write_bytes(b"\0" * 400)
    with pytest.raises(GFError, match="not a LIME file"):
        ildg.read_ildg(tmp_path / "junk.ildg", (2, 2, 2, 2))
    with pytest.raises(ValueError):
        ildg.write_ildg(tmp_path / "bad.ildg", (2, 2, 2, 2), U[:, :8], 64)
    # an unwritable destination (the reference broadcasts such root-rank I/O failures, API.jl:545-561)
    with pytest.raises(GFError, match="cannot"):
        ildg.write_ildg(tmp_path / "no_such_dir" / "c.ildg", (2, 2, 2, 2), U, 64)


def test_api_header(gf):
    from gaugefields_jl_amd import api
    h = api.ILDG(FIXTURE)
    assert len(h) == 1 and h[1]["NC"] == 3 and h[1]["precision"] == 64 and h[1]["L"] is None
    with pytest.raises(IndexError):
        h[2]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_gpu_load_fixture_plaquette(gf, ildg, oracle, precision, tmp_path):
    U, _ = ildg.read_ildg(FIXTURE, DIMS)
    ref = oracle.plaquette_sum(DIMS, U)
    with gf.Lattice(DIMS, precision) as L:
        assert L.load_ildg(FIXTURE) == 64
        tol = 1e-12 if precision == "f64" else 1e-6
        assert L.plaquette_sum() == pytest.approx(ref, rel=tol)
        out = tmp_path / "gpu.ildg"
        L.save_ildg(out, 64)
        U2, _ = ildg.read_ildg(out, DIMS)
        # rows 0 and 1 are stored, row 2 is rebuilt as conj(row0 x row1) on the device
        M, M2 = U.reshape(4, -1, 3, 3).swapaxes(-1, -2), U2.reshape(4, -1, 3, 3).swapaxes(-1, -2)
        if precision == "f64":
            assert np.array_equal(M2[..., :2, :], M[..., :2, :])
            assert np.abs(M2[..., 2, :] - M[..., 2, :]).max() < 1e-14
        else:
            M32 = M.astype(np.complex64).astype(np.complex128)
            assert np.array_equal(M2[..., :2, :], M32[..., :2, :])
            assert np.abs(M2[..., 2, :] - M[..., 2, :]).max() < 1e-6


@pytest.mark.gpu
def test_gpu_api_save_load_configuration(gf, tmp_path):
    from gaugefields_jl_amd import api
    U = api.gauge_configuration(DIMS, start="hot", seed=5)
    p0 = api.measure_plaquette(U)
    path = tmp_path / "conf.ildg"
    api.save_configuration(path, U, format="ildg")
    with pytest.raises(api.ArgumentError):
        api.save_binarydata(U, tmp_path / "x.ildg", precision=16)
    with pytest.raises(api.ArgumentError):
        api.save_configuration(path, U, format="jld2")
    V = api.gauge_configuration(DIMS, start="cold")
    h = api.ILDG(path)
    assert h[1]["L"] == list(DIMS) and h[1]["precision"] == 64
    api.load_gaugefield(V, 1, h, list(DIMS), 3)
    assert api.measure_plaquette(V) == p0
    W = api.gauge_configuration(DIMS, start="cold")
    api.load_configuration_(W, path, format="ildg")
    assert np.array_equal(W.to_host(), U.to_host())
