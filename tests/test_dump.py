"""RCDUMP01 files (rc_dump_*, host only): SHA-256 against hashlib (a
library) at the block-padding edges, round trips, and every corruption the
reader must refuse."""
import hashlib
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    return rc


@pytest.mark.parametrize("n", [0, 1, 3, 55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 1000, 100003])
def test_sha256_matches_hashlib(rc, n):
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    assert rc.sha256(data) == hashlib.sha256(data).hexdigest()


def test_sha256_known_answers(rc):
    # FIPS 180-4 examples
    assert rc.sha256(b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert rc.sha256(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert rc.sha256(b"a" * 1000000) == hashlib.sha256(b"a" * 1000000).hexdigest()


def _planes():
    rng = np.random.default_rng(1)
    return {"depth": rng.random((37, 53), np.float32),
            "emissive": rng.random((3, 37, 53), np.float32),
            "A2": rng.random((16 * 4, 5, 7), np.float32),
            "empty": np.zeros((1, 0, 4), np.float32)}


def test_round_trip(rc, tmp_path):
    p = str(tmp_path / "f.rcdump")
    planes = _planes()
    rc.dump_write(p, planes)
    info = rc.dump_info(p)
    assert info == {"depth": (1, 37, 53), "emissive": (3, 37, 53), "A2": (64, 5, 7), "empty": (1, 0, 4)}
    back = rc.dump_read(p)
    for k, v in planes.items():
        assert np.array_equal(back[k].reshape(v.shape), v)
    # the header's checksum is the SHA-256 of the plane records and the
    # payload (the planes' bytes in order)
    raw = open(p, "rb").read()
    payload = b"".join(np.ascontiguousarray(v, np.float32).tobytes() for v in planes.values())
    assert raw[:8] == b"RCDUMP01" and raw[24:56] == hashlib.sha256(raw[56:]).digest()
    assert raw[-len(payload):] == payload and len(raw) == 56 + 48 * 4 + len(payload)


def test_rejects_corruption(rc, tmp_path):
    p = str(tmp_path / "f.rcdump")
    rc.dump_write(p, _planes())
    raw = bytearray(open(p, "rb").read())

    def bad(mutated, what):
        q = str(tmp_path / "bad.rcdump")
        open(q, "wb").write(bytes(mutated))
        with pytest.raises(rc.RcError, match=what):
            rc.dump_read(q)

    flip = bytearray(raw)
    flip[-5] ^= 0x01
    bad(flip, "checksum")
    bad(raw[:-3], "truncated")
    bad(raw[:40], "truncated header")
    bad(raw + b"x", "trailing")
    m = bytearray(raw)
    m[0:8] = b"RCDUMP02"
    bad(m, "not an RCDUMP01")
    v = bytearray(raw)
    v[8] = 2
    bad(v, "version")
    r = bytearray(raw)
    r[56 + 16] ^= 0x01          # channels of the first plane
    bad(r, "checksum")
    nm = bytearray(raw)
    nm[56] ^= 0x20              # the first plane's name
    bad(nm, "checksum")
    with pytest.raises(rc.RcError, match="cannot open"):
        rc.dump_read(str(tmp_path / "missing.rcdump"))


def test_write_rejects_bad_planes(rc, tmp_path):
    with pytest.raises(ValueError):
        rc.dump_write(str(tmp_path / "x"), {"a" * 17: np.zeros((1, 1), np.float32)})
    with pytest.raises(rc.RcError, match="duplicate"):
        import ctypes
        from radiancecascade_amd import _abi
        a = np.zeros((1, 1, 1), np.float32)
        recs = (_abi.Plane * 2)()
        for i in range(2):
            recs[i].name, recs[i].channels, recs[i].height, recs[i].width, recs[i].data = b"same", 1, 1, 1, a.ctypes.data
        from radiancecascade_amd.stage import _dump_check
        _dump_check(_abi.load().rc_dump_write(os.fsencode(str(tmp_path / "y")), recs, 2))


def test_training_rows_need_matching_gbuffer_planes(rc, tmp_path):
    # rows_from_dumps checks every file's planes before it touches a device
    from radiancecascade_amd.train import rows_from_dumps
    p = _planes()
    no_emis = str(tmp_path / "no_emis.rcdump")
    rc.dump_write(no_emis, {"depth": p["depth"]})
    with pytest.raises(ValueError, match="needs planes"):
        rows_from_dumps([no_emis], levels=6)
    mismatch = str(tmp_path / "mismatch.rcdump")
    rc.dump_write(mismatch, {"depth": p["depth"], "emissive": p["emissive"][:, :, :50]})
    with pytest.raises(ValueError, match="same size"):
        rows_from_dumps([mismatch], levels=6)
    two_ch = str(tmp_path / "two_ch.rcdump")
    rc.dump_write(two_ch, {"depth": np.stack([p["depth"]] * 2), "emissive": p["emissive"]})
    with pytest.raises(ValueError, match="same size"):
        rows_from_dumps([two_ch], levels=6)
