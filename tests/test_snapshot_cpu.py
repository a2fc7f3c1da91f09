"""The state snapshot format (emsph_snapshot_*; no counterpart in the
reference) on the CPU: exact round trips including NaN and signed-zero bit
patterns, the header's fields at their documented offsets, the payload hash
(FNV-1a 64, checked against a direct computation), and rejection of a
corrupted, truncated or foreign buffer."""

import struct

import numpy as np
import pytest

from emerald_amd import _capi
from emerald_amd import sph2d_box as sb


def _fnv1a(b: bytes) -> int:
    h = 0xcbf29ce484222325
    for x in b:
        h ^= x
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _state(n=37, seed=3, colors=True):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 2), dtype=np.float32)
    vel = rng.standard_normal((n, 2)).astype(np.float32)
    pos[0, 0] = np.float32(np.nan)
    vel[1, 1] = np.float32(-0.0)
    col = rng.random((n, 4), dtype=np.float32) if colors else None
    return sb.State(pos, vel, col)


@pytest.mark.parametrize("colors", [True, False])
def test_round_trip_is_exact(colors):
    st = _state(colors=colors)
    data = sb.encode_snapshot(st, time_flicks=-987654321012, method="dfsph_p")
    back, t, m = sb.decode_snapshot(data)
    assert t == -987654321012 and m == "dfsph_p"
    assert np.array_equal(back.positions.view(np.uint32), st.positions.view(np.uint32))
    assert np.array_equal(back.velocities.view(np.uint32), st.velocities.view(np.uint32))
    if colors:
        assert np.array_equal(back.colors.view(np.uint32), st.colors.view(np.uint32))
    else:
        assert back.colors is None


def test_header_layout_and_hash():
    st = _state(n=5)
    data = sb.encode_snapshot(st, time_flicks=23520000, method="iisph")
    assert len(data) == 48 + 5 * 4 * (2 + 2 + 4)
    magic, ver, flags, n, t, m, res, h = struct.unpack_from("<8sIIQqiIQ", data, 0)
    assert magic == b"EMSPHST\x00" and ver == 1 and flags == 3 and n == 5
    assert t == 23520000 and m == _capi.IISPH and res == 0
    assert h == _fnv1a(data[48:])
    payload = np.frombuffer(data[48:], dtype="<f4")
    assert np.array_equal(payload[:10].view(np.uint32), st.positions.ravel().view(np.uint32))
    assert np.array_equal(payload[10:20].view(np.uint32), st.velocities.ravel().view(np.uint32))


def test_empty_state():
    st = sb.State(np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32))
    back, t, m = sb.decode_snapshot(sb.encode_snapshot(st, 7))
    assert len(back.positions) == 0 and t == 7 and m is None


def test_corruption_is_rejected():
    data = bytearray(sb.encode_snapshot(_state(n=9), 1, "pcisph"))
    bad = bytearray(data)
    bad[60] ^= 0x01  # one payload bit
    with pytest.raises(_capi.EmsphError, match="checksum"):
        sb.decode_snapshot(bytes(bad))
    with pytest.raises(_capi.EmsphError, match="truncated"):
        sb.decode_snapshot(bytes(data[:-4]))
    foreign = bytearray(data)
    foreign[0:8] = b"NOTASNAP"
    with pytest.raises(_capi.EmsphError, match="not a snapshot"):
        sb.decode_snapshot(bytes(foreign))
    newer = bytearray(data)
    newer[8] = 2
    with pytest.raises(_capi.EmsphError, match="version"):
        sb.decode_snapshot(bytes(newer))


def test_file_round_trip(tmp_path):
    st = _state(n=11)
    p = str(tmp_path / "s.emsph")
    sb.save_state(p, st, 42, "iisph_pseudo_ap")
    back, t, m = sb.load_state(p)
    assert t == 42 and m == "iisph_pseudo_ap"
    assert np.array_equal(back.velocities.view(np.uint32), st.velocities.view(np.uint32))
