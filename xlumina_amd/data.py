"""Training-pair files and discovered-parameter files for the optimisers.

MultiHDF5DataLoader mirrors xlumina/toolbox.py:698-739: every call picks one
file of a directory at random, then `batch_size` sorted random sample indices
of it, and returns the ("Input fields", "Target fields") rows of those
indices (the datasets generate_synthetic_data.py:102-105 writes).  The rows
are read into pinned host memory and copied to the GPU asynchronously on the
current stream, so a batch is resident in HBM when the next launch reads it.

File formats:
  * ``.npz`` with arrays "Input fields" and "Target fields" (read with
    numpy.load, allow_pickle=False; save_training_pairs writes one);
  * ``.hdf5`` with the same two datasets, when h5py is importable (it is an
    optional dependency; without it .hdf5 files raise ImportError on use).

Discovered parameters (the reference stores {"Discovered parameters": [...]}
with np.save of a pickled dict, examples/noisy_4f_system.ipynb:463-466) are
saved here as a .npz of named arrays "Discovered parameters/<i>", which loads
without unpickling.
"""
from __future__ import annotations

import os
import random
from typing import List, Optional, Sequence

import numpy as np
import torch

KEY_IN, KEY_OUT = "Input fields", "Target fields"
PARAM_KEY = "Discovered parameters"


def _h5py():
    try:
        import h5py  # noqa: F401
    except ImportError as e:   # optional dependency
        raise ImportError("reading .hdf5 training files needs h5py, which is not installed; "
                          "convert them to .npz (save_training_pairs) or install h5py") from e
    return h5py


def save_training_pairs(path: str, inputs, targets) -> None:
    """Writes one .npz training file with the reference's dataset names."""
    inputs, targets = np.asarray(inputs), np.asarray(targets)
    if inputs.shape[0] != targets.shape[0]:
        raise ValueError("inputs and targets must hold the same number of samples")
    np.savez(path, **{KEY_IN: inputs, KEY_OUT: targets})


class MultiHDF5DataLoader:
    """toolbox.py:698-739 with batches delivered as GPU tensors."""

    def __init__(self, directory: str, batch_size: int, device=None, seed: Optional[int] = None,
                 dtype: torch.dtype = torch.complex64):
        self.directory = directory
        self.batch_size = batch_size
        self.files = sorted(f for f in os.listdir(directory) if f.endswith((".hdf5", ".npz")))
        if not self.files:
            raise ValueError(f"No HDF5 files found in directory: {directory}")
        self.device = torch.device(device) if device is not None else (
            torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu"))
        self.dtype = dtype
        self.rng = random.Random(seed)
        self.last_file: Optional[str] = None
        self.last_indices: Optional[List[int]] = None

    def __iter__(self):
        return self

    @staticmethod
    def _read(path: str, idx: Sequence[int]):
        if path.endswith(".npz"):
            with np.load(path, allow_pickle=False) as z:
                return z[KEY_IN][list(idx)], z[KEY_OUT][list(idx)]
        h5 = _h5py()
        with h5.File(path, "r") as hf:
            return hf[KEY_IN][list(idx)], hf[KEY_OUT][list(idx)]

    def _total(self, path: str) -> int:
        if path.endswith(".npz"):
            with np.load(path, allow_pickle=False) as z:
                return int(z[KEY_IN].shape[0])
        h5 = _h5py()
        with h5.File(path, "r") as hf:
            return int(hf[KEY_IN].shape[0])

    def _to_device(self, a: np.ndarray) -> torch.Tensor:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(self.dtype)
        if self.device.type == "cuda":
            t = t.pin_memory().to(self.device, non_blocking=True)
        return t

    def __next__(self):
        """One random file, batch_size sorted random indices -> (inputs, targets)."""
        name = self.rng.choice(self.files)
        path = os.path.join(self.directory, name)
        total = self._total(path)
        idx = sorted(self.rng.sample(range(total), self.batch_size))
        xin, xout = self._read(path, idx)
        self.last_file, self.last_indices = name, idx
        return self._to_device(xin), self._to_device(xout)


def save_parameters(path: str, params: Sequence) -> None:
    """Saves a list of parameter arrays (or scalars) as .npz entries
    "Discovered parameters/<i>"."""
    arrays = {f"{PARAM_KEY}/{i}": (p.detach().cpu().numpy() if isinstance(p, torch.Tensor) else np.asarray(p))
              for i, p in enumerate(params)}
    arrays[f"{PARAM_KEY}/count"] = np.asarray(len(params))
    np.savez(path, **arrays)


def load_parameters(path: str, device=None) -> List[torch.Tensor]:
    """Loads what save_parameters wrote (no unpickling), as tensors on `device`
    (CPU when None)."""
    with np.load(path, allow_pickle=False) as z:
        n = int(z[f"{PARAM_KEY}/count"])
        out = [torch.from_numpy(np.array(z[f"{PARAM_KEY}/{i}"])) for i in range(n)]
    return [t.to(device) for t in out] if device is not None else out
