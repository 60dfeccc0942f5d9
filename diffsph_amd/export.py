"""Frame export in the reference's simulation-export layout, off the hot path.

Drop-ins for ``createOutputFile(config)`` and ``writeFrame(simulationDataGroup,
perennialState, priorState, frameStatistics, config)`` (src/diffSPH/v2/io.py:
8-83): the same file attributes, ``config/<section>`` attributes and one
``simulationExport/%05d`` group per frame holding ``fluidPosition``,
``fluidVelocity``, ``fluidDensity`` (divided by rho0), ``fluidGravity`` (with
gravity), ``fluidShiftAmount``, ``UID`` and the ``boundary*`` datasets, with
the frame statistics as the group's attributes.

The container is a directory tree that mirrors the HDF5 groups
(``<exportPath>/<name>.hdf5.d/``: a group is a directory with an
``attrs.json``; a dataset is a ``.npy`` file), because h5py is not a
dependency here; :func:`readFrame` reads a frame back.

``writeFrame`` never synchronises the simulation: each field is snapshotted
on the device (on the caller's stream), copied device-to-host into pinned
buffers on a side stream and written by a writer thread that waits for the
copies, so the next steps overlap the transfer and the disk.
``close()`` (or leaving a ``with`` block) waits for every queued frame.
"""
from __future__ import annotations

import json
import os
import queue
import threading

import numpy as np
import torch


def _jsonable(v):
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().numpy()
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, (np.floating, np.integer, np.bool_)):
        return v.item()
    if isinstance(v, (bool, int, float, str)) or v is None:
        return v
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    return None


def _scalar(v):
    if isinstance(v, torch.Tensor):
        return float(v.detach().cpu().reshape(-1)[0])
    return float(v)


class Group:
    """One group of the export tree: attributes plus child groups and
    datasets (written by the owning :class:`ExportFile`)."""

    def __init__(self, owner: "ExportFile", path: str):
        self.owner = owner
        self.path = path
        self.attrs = _Attrs(self)
        os.makedirs(path, exist_ok=True)

    def create_group(self, name: str) -> "Group":
        return Group(self.owner, os.path.join(self.path, name))

    def create_dataset(self, name: str, data):
        self.owner._enqueue(os.path.join(self.path, name + ".npy"), data)


class _Attrs(dict):
    def __init__(self, group: Group):
        super().__init__()
        self._g = group

    def __setitem__(self, k, v):
        super().__setitem__(k, _jsonable(v))
        self._g.owner._write_attrs(self._g.path, dict(self))


class ExportFile(Group):
    """The export container with its asynchronous writer."""

    def __init__(self, path: str):
        self._q: queue.Queue = queue.Queue()
        self._err = None
        self._stream = None
        self._thread = threading.Thread(target=self._run, name="diffsph-export", daemon=True)
        self._thread.start()
        super().__init__(self, path)

    # -- writer ------------------------------------------------------------
    def _run(self):
        while True:
            job = self._q.get()
            if job is None:
                self._q.task_done()
                return
            try:
                kind = job[0]
                if kind == "attrs":
                    _, path, attrs = job
                    with open(os.path.join(path, "attrs.json"), "w") as f:
                        json.dump(attrs, f)
                else:
                    _, path, host, event, _keep = job
                    if event is not None:
                        event.synchronize()
                    np.save(path, host.numpy() if isinstance(host, torch.Tensor) else host)
            except Exception as e:  # reported by close()
                self._err = e
            finally:
                self._q.task_done()

    def _write_attrs(self, path, attrs):
        self._q.put(("attrs", path, attrs))

    def _enqueue(self, path, data):
        if isinstance(data, torch.Tensor) and data.is_cuda:
            dev = data.device
            if self._stream is None:
                self._stream = torch.cuda.Stream(device=dev)
            # a device-side snapshot on the caller's stream (microseconds),
            # so the caller may overwrite the field right away; the
            # device-to-host copy of the snapshot runs on the side stream
            src = data.detach().clone()
            host = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
            self._stream.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(self._stream):
                host.copy_(src, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self._stream)
            src.record_stream(self._stream)
            # the snapshot stays referenced until the copy is done
            self._q.put(("data", path, host, ev, src))
        else:
            arr = data.detach().cpu().numpy().copy() if isinstance(data, torch.Tensor) else np.array(data)
            self._q.put(("data", path, arr, None, None))

    def flush(self):
        self._q.join()
        if self._err is not None:
            e, self._err = self._err, None
            raise RuntimeError(f"frame export failed: {e}") from e

    def close(self):
        self.flush()
        self._q.put(None)
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def createOutputFile(config):
    """io.py:8-53: the container, its attributes and config groups; returns
    (file, simulationExport group)."""
    plot = config.get("plot", {})
    name = plot.get("namingScheme", "timestep")
    if name == "timestep":
        name = config["simulation"]["timestamp"]
    os.makedirs(config["export"]["exportPath"], exist_ok=True)
    f = ExportFile(os.path.join(config["export"]["exportPath"], f"{name}.hdf5.d"))
    f.attrs["support"] = _scalar(config["particle"]["support"])
    f.attrs["targetNeighbors"] = _scalar(config["kernel"]["targetNeighbors"])
    f.attrs["restDensity"] = config["fluid"]["rho0"]
    f.attrs["radius"] = _scalar(config["particle"]["dx"])
    f.attrs["area"] = _scalar(config["particle"]["dx"]) ** 2
    if not config.get("gravity", {}).get("active", False):
        f.attrs["fluidGravity"] = [0, 0]
    cg = f.create_group("config")
    for key, section in config.items():
        if not isinstance(section, dict):
            continue
        g = cg.create_group(key)
        if key == "regions":
            continue
        for sub, val in section.items():
            if isinstance(val, (dict, torch.dtype, torch.device)) or callable(val):
                continue
            j = _jsonable(val)
            if j is not None:
                g.attrs[sub] = j
    return f, f.create_group("simulationExport")


def writeFrame(simulationDataGroup, perennialState, priorState, frameStatistics, config):
    """io.py:55-83: one frame group named by the time step."""
    g = simulationDataGroup.create_group(f'{perennialState["timestep"]:05d}')
    for k, v in frameStatistics.items():
        g.attrs[k] = v
    rho0 = config["fluid"]["rho0"]
    fl = priorState["fluid"]
    g.create_dataset("fluidPosition", fl["positions"])
    g.create_dataset("fluidVelocity", fl["velocities"])
    g.create_dataset("fluidDensity", fl["densities"] / rho0)
    if config.get("gravity", {}).get("active", False):
        g.create_dataset("fluidGravity", fl["gravityAccel"])
    g.create_dataset("fluidShiftAmount", perennialState["fluid"]["shiftAmount"])
    g.create_dataset("UID", perennialState["fluid"]["index"])
    if "boundary" in perennialState and perennialState["boundary"] is not None:
        b = priorState["boundary"]
        g.create_dataset("boundaryPosition", b["positions"])
        g.create_dataset("boundaryVelocity", b["velocities"])
        g.create_dataset("boundaryDensity", b["densities"] / rho0)
        g.create_dataset("boundaryUID", b["index"])
    return g


def readFrame(path: str, timestep: int) -> dict:
    """Datasets and attributes of one written frame (for checks and
    post-processing)."""
    d = os.path.join(path, "simulationExport", f"{timestep:05d}")
    out = {}
    for fn in sorted(os.listdir(d)):
        if fn.endswith(".npy"):
            out[fn[:-4]] = np.load(os.path.join(d, fn), allow_pickle=False)
    ap = os.path.join(d, "attrs.json")
    if os.path.exists(ap):
        with open(ap) as f:
            out["attrs"] = json.load(f)
    return out
