"""Training-pair loader and parameter files (CPU; the GPU copy is covered by
test_gpu_api's loader test)."""
import os

import numpy as np
import pytest
import torch

from xlumina_amd import data as D


def test_loader_picks_sorted_indices_of_one_file(tmp_path):
    rng = np.random.default_rng(0)
    files = {}
    for k in range(3):
        xin = rng.standard_normal((9, 6, 6)) + 1j * rng.standard_normal((9, 6, 6))
        xout = rng.standard_normal((9, 6, 6))
        D.save_training_pairs(str(tmp_path / f"set_{k}.npz"), xin, xout)
        files[f"set_{k}.npz"] = (xin, xout)
    L = D.MultiHDF5DataLoader(str(tmp_path), batch_size=4, device="cpu", seed=3)
    for _ in range(5):
        a, b = next(L)
        xin, xout = files[L.last_file]
        assert L.last_indices == sorted(set(L.last_indices)) and len(L.last_indices) == 4
        assert a.dtype == torch.complex64 and a.shape == (4, 6, 6)
        np.testing.assert_allclose(a.numpy(), xin[L.last_indices].astype(np.complex64))
        np.testing.assert_allclose(b.numpy().real, xout[L.last_indices].astype(np.float32))


def test_loader_errors(tmp_path):
    with pytest.raises(ValueError):
        D.MultiHDF5DataLoader(str(tmp_path), 2)
    D.save_training_pairs(str(tmp_path / "a.npz"), np.zeros((2, 3, 3)), np.zeros((2, 3, 3)))
    with pytest.raises(ValueError):
        next(D.MultiHDF5DataLoader(str(tmp_path), 5, device="cpu", seed=0))
    with pytest.raises(ValueError):
        D.save_training_pairs(str(tmp_path / "b.npz"), np.zeros((2, 3, 3)), np.zeros((3, 3, 3)))
    (tmp_path / "c.hdf5").write_bytes(b"")
    try:
        import h5py  # noqa: F401
    except ImportError:
        L = D.MultiHDF5DataLoader(str(tmp_path), 1, device="cpu", seed=0)
        L.files = ["c.hdf5"]
        with pytest.raises(ImportError):
            next(L)


def test_parameters_round_trip(tmp_path):
    params = [np.arange(6.0).reshape(2, 3), torch.tensor([0.5, -1.0]), 2.25]
    p = str(tmp_path / "found.npz")
    D.save_parameters(p, params)
    back = D.load_parameters(p)
    assert len(back) == 3
    np.testing.assert_array_equal(back[0].numpy(), params[0])
    np.testing.assert_array_equal(back[1].numpy(), params[1].numpy())
    assert float(back[2]) == 2.25
    with np.load(p, allow_pickle=False) as z:   # plain arrays, nothing pickled
        assert all(z[k].dtype != object for k in z.files)
