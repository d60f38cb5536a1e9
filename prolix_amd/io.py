"""Simulation-state records and trajectory files.

Mirrors prolix's ``SimulationState`` (prolix/typing.py:653-728) and
``TrajectoryWriter`` (prolix/simulate.py:148-203):

- a state is positions, velocities, step and time (ns), optionally forces,
  masses and the potential / kinetic energies;
- ``to_array_record`` packs ``numpy()`` minus its None fields with msgpack.
  Arrays and numpy scalars use the wire format of msgpack-numpy 0.4.8 (the
  reference's pinned encoder, not installed here), restated in
  ``_encode`` / ``_decode``: an array becomes the map
  {b"nd": True, b"type": dtype.str, b"kind": b"", b"shape": shape, b"data": bytes},
  and a numpy scalar becomes {b"nd": False, b"type": dtype.str, b"data": bytes}.
  Records written here therefore unpack with msgpack-numpy, and theirs with
  ``from_array_record``;
- ArrayRecord (the reference's container, ``array_record``) is not
  installed either. ``TrajectoryWriter`` stores the same per-frame records
  in a length-prefixed file: the magic ``PRXTRJ1\\n``, then per record a
  little-endian uint64 length and the bytes.

Velocities of engine states are in A per AKMA time unit (48.88821291839 fs),
the units the engine's integrator runs in.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import Any, Iterator, Sequence

import msgpack
import numpy as np

MAGIC = b"PRXTRJ1\n"


def _encode(obj):
    if isinstance(obj, np.ndarray):
        if obj.dtype.kind == "V":
            raise TypeError("structured arrays are not part of simulation states")
        return {b"nd": True, b"type": obj.dtype.str, b"kind": b"", b"shape": list(obj.shape),
                b"data": np.ascontiguousarray(obj).tobytes()}
    if isinstance(obj, (np.bool_, np.number)):
        return {b"nd": False, b"type": obj.dtype.str, b"data": obj.tobytes()}
    return obj


def _decode(obj):
    if b"nd" in obj:
        if obj[b"nd"] is True:
            if obj.get(b"kind") == b"V":
                raise TypeError("structured arrays are not part of simulation states")
            return np.ndarray(buffer=obj[b"data"], dtype=np.dtype(obj[b"type"]), shape=tuple(obj[b"shape"])).copy()
        return np.frombuffer(obj[b"data"], dtype=np.dtype(obj[b"type"]))[0]
    if b"complex" in obj:
        data = obj[b"data"]
        return complex(data.decode() if isinstance(data, bytes) else data)
    return obj


def packb(data: Any) -> bytes:
    return msgpack.packb(data, default=_encode, use_bin_type=True)


def unpackb(packed: bytes) -> Any:
    return msgpack.unpackb(packed, object_hook=_decode, raw=False)


@dataclass
class SimulationState:
    """One frame (prolix/typing.py:653-728)."""

    positions: Any
    velocities: Any
    step: Any
    time_ns: Any
    forces: Any = None
    mass: Any = None
    potential_energy: Any = None
    kinetic_energy: Any = None
    # The integrator's own float32 momenta (amu A per AKMA time unit), beside
    # the velocities: with ``step`` they resume a trajectory bit for bit. An
    # extra record key that prolix's reader (typing.py:706-728) skips.
    momenta: Any = None

    @property
    def position(self):
        return self.positions

    def numpy(self) -> dict:
        def cpu(x):
            if x is None or isinstance(x, (int, float)):
                return x
            if hasattr(x, "detach"):
                x = x.detach().cpu().numpy()
            return np.asarray(x)

        return {k: cpu(getattr(self, k)) for k in ("positions", "velocities", "forces", "mass", "step", "time_ns",
                                                     "potential_energy", "kinetic_energy", "momenta")}

    def to_array_record(self) -> bytes:
        return packb({k: v for k, v in self.numpy().items() if v is not None})

    @classmethod
    def from_array_record(cls, packed: bytes) -> "SimulationState":
        d = unpackb(packed)
        return cls(positions=d["positions"], velocities=d["velocities"], step=d["step"], time_ns=d["time_ns"],
                   forces=d.get("forces"), mass=d.get("mass"), potential_energy=d.get("potential_energy"),
                   kinetic_energy=d.get("kinetic_energy"), momenta=d.get("momenta"))


class TrajectoryWriter:
    """Appends per-frame state records to a trajectory file (prolix/simulate.py:148-203)."""

    def __init__(self, path: str):
        self._fh = open(path, "wb")
        self._fh.write(MAGIC)
        self.closed = False

    def _put(self, rec: bytes):
        if self.closed:
            raise RuntimeError("Writer is closed")
        self._fh.write(struct.pack("<Q", len(rec)))
        self._fh.write(rec)

    def write(self, states: SimulationState | Sequence[SimulationState]) -> None:
        if isinstance(states, SimulationState):
            states = [states]
        for st in states:
            self._put(st.to_array_record())

    def write_batch(self, stacked: SimulationState) -> None:
        """One record per leading index of a stacked state (scalars are shared)."""
        data = stacked.numpy()
        n = np.asarray(data["positions"]).shape[0]
        for i in range(n):
            frame = {k: (v[i] if v is not None and getattr(v, "ndim", 0) > 0 else v) for k, v in data.items()
                     if v is not None}
            self._put(packb(frame))

    def close(self) -> None:
        if not self.closed:
            self._fh.close()
            self.closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if not getattr(self, "closed", True):
            self.close()


def read_trajectory(path: str) -> Iterator[SimulationState]:
    """Frames of a file written by TrajectoryWriter, in order."""
    with open(path, "rb") as fh:
        if fh.read(len(MAGIC)) != MAGIC:
            raise ValueError(f"{path}: not a trajectory file")
        while True:
            head = fh.read(8)
            if not head:
                return
            if len(head) != 8:
                raise ValueError(f"{path}: truncated record header")
            (n,) = struct.unpack("<Q", head)
            rec = fh.read(n)
            if len(rec) != n:
                raise ValueError(f"{path}: truncated record")
            yield SimulationState.from_array_record(rec)


def engine_states(engine, masses_amu, dt_fs: float, energies: bool = True) -> list[SimulationState]:
    """One state per system of a GBEngine, read from its device-resident integrator state.

    ``masses_amu`` are the packed per-atom masses the integrator runs with;
    velocities are momentum / mass.
    """
    m = np.asarray(masses_amu, np.float64).reshape(-1)
    if m.shape[0] != engine.n_atoms:
        raise ValueError("one mass per packed atom is required")
    x, p, f = engine.state()
    xs, ps, fs = engine.unpack(x), engine.unpack(p), engine.unpack(f)
    e = engine.energy_forces(x)[0].cpu().numpy() if energies else None
    ke = engine.kinetic().cpu().numpy() if energies else None
    out = []
    for k in range(engine.n_systems):
        mk = m[engine.offsets[k]:engine.offsets[k + 1]]
        steps = engine.step_count(k)
        out.append(SimulationState(
            positions=np.asarray(xs[k], np.float32), velocities=(ps[k] / mk[:, None]).astype(np.float32),
            step=np.int64(steps), time_ns=np.float64(steps * dt_fs * 1e-6), forces=np.asarray(fs[k], np.float32),
            mass=mk, potential_energy=None if e is None else np.float64(e[k, -1]),
            kinetic_energy=None if ke is None else np.float64(ke[k]), momenta=np.asarray(ps[k], np.float32)))
    return out


def restore_engine(engine, states, params) -> None:
    """Load one saved state per system into ``engine`` and continue from it.

    Positions, the saved float32 momenta and the step count (the position of
    each system's Philox noise stream) are restored, and the forces are
    recomputed at the saved positions, so the continued trajectory equals an
    uninterrupted one bit for bit (the resumed production of
    batched_simulate.py:1300-1305). ``params``: one Langevin per system (or one
    for all). States written without momenta cannot be resumed exactly and
    raise ValueError.
    """
    if len(states) != engine.n_systems:
        raise ValueError("one state per system is required")
    if any(st.momenta is None for st in states):
        raise ValueError("a state without the integrator's momenta cannot resume a trajectory exactly")
    x = [np.asarray(st.positions, np.float32).reshape(-1, 3) for st in states]
    p = [np.asarray(st.momenta, np.float32).reshape(-1, 3) for st in states]
    engine.init(x, params, momenta=p)
    engine.set_step_counts([int(np.asarray(st.step)) for st in states])


DEFAULT_RECENTER_EVERY = 100  # steps between centre-of-mass re-centrings (batched_simulate.py:55)


def _next_recenter(step: int, every: int) -> int:
    """Step count after which the next re-centring falls: the step with 0-based
    index i > 0 and i % every == 0 ends at count i + 1."""
    j = max(1, (step - 1) // every + 1)
    return j * every + 1


def produce_streaming(engine, n_saves: int, steps_per_save: int, write_fn, *,
                      recenter_every: int = DEFAULT_RECENTER_EVERY, write_batch_size: int = 1,
                      start_frame: int = 0, device_offset: int = 0) -> int:
    """Streaming production on an initialised GBEngine (batched_simulate.py:1226-1422).

    Runs (n_saves - start_frame) * steps_per_save steps of the engine's BAOAB
    integrator and returns that count. Nothing accumulates on the device:

    - after the step with 0-based index i (counted from this call) where
      i > 0 and i % recenter_every == 0, every system's positions are
      re-centred on its mass-weighted centre of mass on the device
      (apply_com_correction without rotation removal, batched_simulate.py:122-155);
      recenter_every <= 0 turns this off;
    - after every steps_per_save steps the positions are copied to pinned host
      memory behind the step launches, and the host writes frame k - 1 while
      the GPU runs the steps of frame k;
    - write_fn(positions (N_s, 3) float32, batch_idx, save_idx) is called per
      system s, with batch_idx = s + device_offset; with write_batch_size
      W > 1 frames are grouped, write_fn(frames (W, N_s, 3), batch_idx,
      start_save_idx), and W must divide n_saves - start_frame.

    Save indices start at start_frame, as in the reference's grouped path.
    Systems are written without padding atoms.
    """
    t = engine.torch
    remaining = int(n_saves) - int(start_frame)
    if remaining < 0 or steps_per_save <= 0 or write_batch_size < 1:
        raise ValueError("need n_saves >= start_frame, steps_per_save > 0 and write_batch_size >= 1")
    if remaining % write_batch_size:
        raise ValueError(f"write_batch_size={write_batch_size} must divide the {remaining} remaining saves")
    offsets = [int(o) for o in engine.offsets]
    on_device = t.cuda.is_available()
    host = [t.empty((engine.n_atoms, 3), dtype=t.float32, pin_memory=on_device) for _ in range(2)]
    groups = [[] for _ in range(engine.n_systems)]
    step = 0

    def run_to(target):
        nonlocal step
        while step < target:
            nxt = target
            if recenter_every > 0:
                nxt = min(nxt, _next_recenter(step, recenter_every))
            engine.steps(nxt - step)
            step = nxt
            if recenter_every > 0 and step > 1 and (step - 1) % recenter_every == 0:
                engine.recenter_com()

    def flush(frame):
        ev, buf, save_idx = frame
        if ev is not None:
            ev.synchronize()
        x = buf.numpy()
        for s in range(engine.n_systems):
            xs = x[offsets[s]:offsets[s + 1]].copy()
            if write_batch_size == 1:
                write_fn(xs, s + device_offset, save_idx)
                continue
            groups[s].append(xs)
            if len(groups[s]) == write_batch_size:
                write_fn(np.stack(groups[s]), s + device_offset, save_idx - write_batch_size + 1)
                groups[s] = []

    pending = None
    for k in range(remaining):
        run_to((k + 1) * steps_per_save)
        x = engine.state()[0]
        buf = host[k % 2]
        buf.copy_(x, non_blocking=on_device)
        ev = None
        if on_device:
            ev = t.cuda.Event()
            ev.record()
        if pending is not None:
            flush(pending)
        pending = (ev, buf, start_frame + k)
    if pending is not None:
        flush(pending)
    return step


def trajectory_write_fn(writers, dt_fs: float, steps_per_save: int):
    """A write_fn for produce_streaming that appends SimulationState frames
    (positions only; velocities zero-filled, as the stream carries positions)
    to one TrajectoryWriter per batch index."""

    def write(positions, batch_idx, save_idx):
        frames = positions if positions.ndim == 3 else positions[None]
        for k, xs in enumerate(frames):
            step = (save_idx + k + 1) * steps_per_save
            writers[batch_idx].write(SimulationState(positions=xs, velocities=np.zeros_like(xs),
                                                     step=np.int64(step), time_ns=np.float64(step * dt_fs * 1e-6)))

    return write
