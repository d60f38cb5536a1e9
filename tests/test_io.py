"""State records and trajectory files (prolix/typing.py:653-728, prolix/simulate.py:148-203).

The wire format is msgpack-numpy 0.4.8's (the reference's pinned encoder,
absent here): the known-answer byte strings below are written out by hand
from the msgpack specification (fixmap / bin8 / fixstr / fixarray / true /
false markers), so they pin the encoding independently of ``prolix_amd.io``.
"""

import struct

import msgpack
import numpy as np
import pytest

from prolix_amd import io

# {"x": np.array([1.5], float32)} and {"s": np.int64(7)} in msgpack-numpy's map form.
ARRAY_KAT = bytes.fromhex(
    "81" "a178"
    "85"
    "c4026e64" "c3"
    "c40474797065" "a33c6634"
    "c4046b696e64" "c400"
    "c40573686170" "65" "9101"
    "c40464617461" "c4040000c03f")
SCALAR_KAT = bytes.fromhex(
    "81" "a173"
    "83"
    "c4026e64" "c2"
    "c40474797065" "a33c6938"
    "c40464617461" "c4080700000000000000")


def test_known_answer_bytes_decode_and_encode():
    d = io.unpackb(ARRAY_KAT)
    assert d["x"].dtype == np.float32 and d["x"].shape == (1,) and d["x"][0] == 1.5
    assert io.packb({"x": np.array([1.5], np.float32)}) == ARRAY_KAT
    s = io.unpackb(SCALAR_KAT)["s"]
    assert isinstance(s, np.int64) and s == 7
    assert io.packb({"s": np.int64(7)}) == SCALAR_KAT


def test_plain_msgpack_sees_the_map_form():
    a = np.arange(6, dtype=np.float64).reshape(2, 3)
    raw = msgpack.unpackb(io.packb({"a": a}), raw=False)["a"]
    assert raw[b"nd"] is True and raw[b"type"] == "<f8" and raw[b"kind"] == b"" and raw[b"shape"] == [2, 3]
    assert np.frombuffer(raw[b"data"], "<f8").tolist() == a.reshape(-1).tolist()
    # Fortran order and non-contiguous views are stored in C order.
    raw2 = msgpack.unpackb(io.packb({"a": np.asfortranarray(a)[:, ::2]}), raw=False)["a"]
    assert raw2[b"shape"] == [2, 2] and np.frombuffer(raw2[b"data"], "<f8").tolist() == [0, 2, 3, 5]


def _state(rng, n=7, full=True):
    return io.SimulationState(
        positions=rng.normal(size=(n, 3)).astype(np.float32), velocities=rng.normal(size=(n, 3)).astype(np.float32),
        step=np.int64(1234), time_ns=np.float64(2.468e-3),
        forces=rng.normal(size=(n, 3)).astype(np.float32) if full else None,
        mass=rng.uniform(1, 16, n) if full else None,
        potential_energy=np.float64(-1234.5) if full else None, kinetic_energy=12.25 if full else None,
        momenta=rng.normal(size=(n, 3)).astype(np.float32) if full else None)


def _same(a, b):
    for k in ("positions", "velocities", "forces", "mass", "step", "time_ns", "potential_energy", "kinetic_energy",
              "momenta"):
        x, y = getattr(a, k), getattr(b, k)
        if x is None:
            assert y is None, k
        else:
            assert np.asarray(x).dtype == np.asarray(y).dtype, k
            assert np.array_equal(np.asarray(x), np.asarray(y)), k


@pytest.mark.parametrize("full", [True, False])
def test_state_round_trip(full):
    st = _state(np.random.default_rng(0), full=full)
    rec = st.to_array_record()
    keys = set(msgpack.unpackb(rec, raw=False))
    assert keys == ({"positions", "velocities", "step", "time_ns"} | (
        {"forces", "mass", "potential_energy", "kinetic_energy", "momenta"} if full else set()))
    back = io.SimulationState.from_array_record(rec)
    _same(st, back)
    assert back.position is back.positions


def test_torch_tensors_are_recorded_as_arrays():
    torch = pytest.importorskip("torch")
    st = io.SimulationState(positions=torch.ones(3, 3), velocities=torch.zeros(3, 3), step=5, time_ns=0.5)
    back = io.SimulationState.from_array_record(st.to_array_record())
    assert back.positions.dtype == np.float32 and back.positions.sum() == 9 and back.step == 5


def test_trajectory_file(tmp_path):
    rng = np.random.default_rng(1)
    frames = [_state(rng, full=k % 2 == 0) for k in range(5)]
    path = tmp_path / "traj.prx"
    with io.TrajectoryWriter(str(path)) as w:
        w.write(frames[0])
        w.write(frames[1:])
    back = list(io.read_trajectory(str(path)))
    assert len(back) == 5
    for a, b in zip(frames, back):
        _same(a, b)
    raw = path.read_bytes()
    assert raw[:8] == io.MAGIC
    (n0,) = struct.unpack("<Q", raw[8:16])
    assert raw[16:16 + n0] == frames[0].to_array_record()
    with pytest.raises(RuntimeError):
        w.write(frames[0])


def test_write_batch_slices_the_leading_axis(tmp_path):
    rng = np.random.default_rng(2)
    b, n = 4, 6
    stacked = io.SimulationState(positions=rng.normal(size=(b, n, 3)), velocities=rng.normal(size=(b, n, 3)),
                                 step=np.arange(b, dtype=np.int64) * 10, time_ns=0.125,
                                 potential_energy=rng.normal(size=b))
    path = tmp_path / "batch.prx"
    w = io.TrajectoryWriter(str(path))
    w.write_batch(stacked)
    w.close()
    back = list(io.read_trajectory(str(path)))
    assert len(back) == b
    for i, st in enumerate(back):
        assert np.array_equal(st.positions, stacked.positions[i])
        assert st.step == 10 * i and st.time_ns == 0.125
        assert st.potential_energy == stacked.potential_energy[i] and st.forces is None


def test_truncated_or_foreign_files_raise(tmp_path):
    path = tmp_path / "t.prx"
    with io.TrajectoryWriter(str(path)) as w:
        w.write(_state(np.random.default_rng(3)))
    raw = path.read_bytes()
    (tmp_path / "cut.prx").write_bytes(raw[:-3])
    with pytest.raises(ValueError, match="truncated"):
        list(io.read_trajectory(str(tmp_path / "cut.prx")))
    (tmp_path / "other.prx").write_bytes(b"NOTATRJ\n" + raw[8:])
    with pytest.raises(ValueError, match="not a trajectory"):
        list(io.read_trajectory(str(tmp_path / "other.prx")))


class _FakeEngine:
    """Host stand-in for GBEngine's streaming surface: each step adds a fixed
    per-system velocity to the positions; re-centring subtracts the mass-weighted
    centre of mass. Records the call sequence."""

    def __init__(self, sizes, seed=0):
        import torch

        self.torch = torch
        self.n_systems = len(sizes)
        self.offsets = np.concatenate([[0], np.cumsum(sizes)])
        self.n_atoms = int(self.offsets[-1])
        rng = np.random.default_rng(seed)
        self.x = rng.normal(size=(self.n_atoms, 3))
        self.m = rng.uniform(1, 16, self.n_atoms)
        self.v = rng.normal(size=(self.n_atoms, 3)) * 0.01
        self.calls = []
        self.count = 0

    def steps(self, n):
        self.calls.append(("steps", n))
        self.x = self.x + n * self.v
        self.count += n

    def recenter_com(self):
        self.calls.append(("recenter", self.count))
        for s in range(self.n_systems):
            sl = slice(self.offsets[s], self.offsets[s + 1])
            self.x[sl] -= (self.m[sl, None] * self.x[sl]).sum(0) / self.m[sl].sum()

    def state(self):
        x = self.torch.tensor(self.x, dtype=self.torch.float32)
        return x, None, None


def test_streaming_schedule_and_frames():
    eng = _FakeEngine([5, 9])
    got = []
    n = io.produce_streaming(eng, 6, 7, lambda x, b, k: got.append((b, k, x)), recenter_every=10)
    assert n == 42 and eng.count == 42
    # re-centrings after the steps with 0-based index 10, 20, 30, 40
    assert [c for name, c in eng.calls if name == "recenter"] == [11, 21, 31, 41]
    assert [(b, k) for b, k, _ in got] == [(b, k) for k in range(6) for b in range(2)]
    # the last frame of system 1: its centre of mass moved one step past the re-centring at 41
    xs = got[-1][2]
    sl = slice(5, 14)
    com = (eng.m[sl, None] * xs).sum(0) / eng.m[sl].sum()
    drift = (eng.m[sl, None] * eng.v[sl]).sum(0) / eng.m[sl].sum()
    assert np.allclose(com, drift, atol=1e-5)


def test_streaming_grouped_writes_and_resume():
    eng = _FakeEngine([4])
    got = []
    io.produce_streaming(eng, 8, 3, lambda x, b, k: got.append((b, k, x.shape)), recenter_every=0,
                         write_batch_size=3, start_frame=2)
    assert got == [(0, 2, (3, 4, 3)), (0, 5, (3, 4, 3))]
    assert all(name == "steps" for name, _ in eng.calls) and eng.count == 18
    with pytest.raises(ValueError, match="must divide"):
        io.produce_streaming(eng, 8, 3, lambda *a: None, write_batch_size=4, start_frame=1)


def test_streaming_into_trajectory_files(tmp_path):
    eng = _FakeEngine([3, 4])
    paths = [tmp_path / f"s{k}.prx" for k in range(2)]
    writers = [io.TrajectoryWriter(str(p)) for p in paths]
    io.produce_streaming(eng, 4, 5, io.trajectory_write_fn(writers, dt_fs=2.0, steps_per_save=5), recenter_every=0)
    for w in writers:
        w.close()
    frames = [list(io.read_trajectory(str(p))) for p in paths]
    assert [len(f) for f in frames] == [4, 4]
    assert [int(st.step) for st in frames[1]] == [5, 10, 15, 20]
    assert frames[1][-1].time_ns == pytest.approx(20 * 2.0e-6)
    assert np.allclose(frames[1][-1].positions, eng.x[3:7], atol=1e-6)


def test_restore_needs_the_integrator_momenta():
    """A record without the float32 momenta (e.g. one written by prolix, whose
    reader and writer know only the velocities) cannot resume bit for bit."""

    class _Eng:
        n_systems = 1

    st = _state(np.random.default_rng(1), full=False)
    with pytest.raises(ValueError, match="momenta"):
        io.restore_engine(_Eng(), [st], None)
    with pytest.raises(ValueError, match="one state per system"):
        io.restore_engine(_Eng(), [st, st], None)
