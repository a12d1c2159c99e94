"""Host-side logic on the CPU: the torch PCG32 / TEA used by the Python
samplers agree with the oracle bit for bit; scene generators give the sizes
the workloads are quoted at."""
import numpy as np
import torch

import oracle as O
from wos import configs as C
from wos.fwd import PCG32
from wos.utils import sample_tea_32


def test_torch_pcg32_matches_oracle():
    s = PCG32(7, 0x853C49E6748FEA9B, 0xDA3E39CB94B95BDB, dev="cpu")
    words = [s.next_uint32().numpy() for _ in range(5)]
    for k in range(7):
        ref = O.pcg32_stream(k, 0x853C49E6748FEA9B, 0xDA3E39CB94B95BDB, 5)
        assert [int(w[k]) for w in words] == ref.tolist()
    f = s.next_float32().numpy()
    d = s.next_float64().numpy()
    assert ((f >= 0) & (f < 1)).all() and ((d >= 0) & (d < 1)).all()


def test_torch_pcg32_tea_seeding_matches_oracle():
    idx = torch.arange(40) // 8
    v0, v1 = sample_tea_32(9, idx)
    for i in range(5):
        assert (int(v0[8 * i]), int(v1[8 * i])) == O.tea(9, i)
    s = PCG32(40, v0.to(torch.int64), v1.to(torch.int64), dev="cpu")
    w = s.next_uint32().numpy()
    for k in (0, 13, 39):
        a, b = O.tea(9, k // 8)
        ref = O.pcg32_stream(k, a & (2 ** 64 - 1), b & (2 ** 64 - 1), 1)
        assert int(w[k]) == int(ref[0])


def test_config_sizes():
    assert C.config1()["points"].shape == (11428, 2)
    assert C.config2()["points"].shape == (126820, 2)
    v, f = C.icosphere(3)
    assert v.shape == (642, 3) and f.shape == (1280, 3)
    n = C.SmoothNoise(0)
    probe, _ = C.icosphere(4)
    assert abs(np.abs(n(probe)).max() - 1.0) < 0.05


def test_autotune_coordinate_descent_finds_separable_minimum():
    from wos_ext.autotune import CANDIDATES, coordinate_descent
    target = (128, 6, 4, 12)

    def measure(t):
        return sum((a - b) ** 2 / (b * b) for a, b in zip(t, target)) + 1.0

    best, ms, trials = coordinate_descent(measure, (64, 8, 0, 4))
    assert best == target and ms == 1.0
    assert trials <= sum(len(v) for v in CANDIDATES.values()) * 2 + 1


def test_autotune_cache_round_trip(tmp_path):
    from wos_ext import autotune as A
    p = str(tmp_path / "sub" / "tuning.json")
    A._store(p, {"dev|k": {"tuning": [128, 8, 6, 8], "ms": 1.5}})
    assert A._load(p)["dev|k"]["tuning"] == [128, 8, 6, 8]
    assert A._load(str(tmp_path / "missing.json")) == {}


def test_exr_round_trip_and_layout(tmp_path):
    """Uncompressed scanline EXR: float and half round trips, and the byte
    layout of a 2x1 single-channel FLOAT file assembled by hand from the
    OpenEXR file-layout specification."""
    import struct
    from wos import io
    rng = np.random.default_rng(0)
    img = rng.normal(size=(5, 7, 3)).astype(np.float32)
    io.write_exr(str(tmp_path / "a.exr"), img)
    np.testing.assert_array_equal(io.read_exr(str(tmp_path / "a.exr")), img)
    io.write_exr(str(tmp_path / "h.exr"), img, half=True)
    np.testing.assert_array_equal(io.read_exr(str(tmp_path / "h.exr")), img.astype(np.float16).astype(np.float32))
    io.write_exr(str(tmp_path / "y.exr"), np.array([[1.5, -2.0]], np.float32))
    b = open(tmp_path / "y.exr", "rb").read()
    assert b[:8] == struct.pack("<ii", 20000630, 2)
    assert b"channels\0chlist\0" + struct.pack("<i", 19) + b"Y\0" + struct.pack("<iB3xii", 2, 0, 1, 1) + b"\0" in b
    assert b"compression\0compression\0" + struct.pack("<i", 1) + b"\0" in b
    assert b"dataWindow\0box2i\0" + struct.pack("<i", 16) + struct.pack("<iiii", 0, 0, 1, 0) in b
    assert b.endswith(struct.pack("<ii", 0, 8) + struct.pack("<ff", 1.5, -2.0))
    off = struct.unpack_from("<Q", b, len(b) - 16 - 8)[0]
    assert off == len(b) - 16


def test_obj_round_trips(tmp_path):
    from wos import io
    v = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    ix = np.array([[0, 1], [1, 2], [2, 0]])
    c = np.array([0.5, -1.0, 2.0])
    io.write_2d_obj(str(tmp_path / "p.obj"), v, ix, c)
    v2, ix2, c2 = io.read_2d_obj(str(tmp_path / "p.obj"))
    np.testing.assert_allclose(v2, v)
    np.testing.assert_array_equal(ix2, ix)
    np.testing.assert_allclose(c2, c)
    _, ixf, _ = io.read_2d_obj(str(tmp_path / "p.obj"), flip_orientation=True)
    np.testing.assert_array_equal(ixf, ix[:, ::-1])
    v3 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    f3 = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    io.write_3d_obj(str(tmp_path / "t.obj"), v3, f3, v3[:, 2])
    a, b, cc = io.read_3d_obj(str(tmp_path / "t.obj"))
    np.testing.assert_allclose(a, v3)
    np.testing.assert_array_equal(b, f3)
    np.testing.assert_allclose(cc, v3[:, 2])


def test_srgb_transfer_is_inverse():
    from wos import io
    x = np.linspace(0, 1, 101)
    np.testing.assert_allclose(io.srgb_to_linear(io.linear_to_srgb(x)), x, atol=1e-12)
    assert abs(io.linear_to_srgb(0.0031308) - 0.0031308 * 12.92) < 1e-9


def test_detector_index_stays_in_the_image():
    """Detector.index (wos/scene.py:18-24) flips rows as res_y - uv.y, which
    sends uv.y = 0 one row past the image: such points must come back
    invalid, and every valid index must address a pixel."""
    import torch
    from wos.scene import Detector
    from wos.scene3d import Detector3D
    det = Detector(vmin=(-1.2, -1.2), vmax=(1.2, 1.2), res=(64, 48))
    rng = np.random.default_rng(0)
    p = torch.from_numpy(rng.uniform(-1.4, 1.4, (200000, 2)).astype(np.float32))
    edge = torch.tensor([[0.0, -1.2], [0.0, -1.2 + 0.01], [-1.2, 0.0], [1.1, 1.1]], dtype=torch.float32)
    valid, idx = det.index(torch.cat([p, edge]))
    n = det.res[0] * det.res[1]
    assert ((idx[valid] >= 0) & (idx[valid] < n)).all()
    assert not bool(valid[-4]) and not bool(valid[-3])     # the bottom row (uv.y = 0)
    assert bool(valid[-2]) and bool(valid[-1])
    # the rows that remain are the reference's: uv.x + (res_y - uv.y) res_x
    uv = ((torch.cat([p, edge]) + 1.2) / 2.4 * torch.tensor([64.0, 48.0]) + 0.5).to(torch.int32)
    ref = uv[:, 0] + (48 - uv[:, 1]) * 64
    assert torch.equal(idx[valid], ref[valid])
    d3 = Detector3D(vmin=(-1.0, -1.0), vmax=(1.0, 1.0), res=(32, 32), z=0.1)
    q = torch.tensor([[0.0, 0.0, 0.1], [0.0, 0.0, 0.5], [0.0, -1.0, 0.1]], dtype=torch.float32)
    v3, i3 = d3.index(q)
    assert v3.tolist() == [True, False, False]
    assert 0 <= int(i3[0]) < 32 * 32
