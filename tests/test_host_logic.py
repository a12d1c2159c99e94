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
