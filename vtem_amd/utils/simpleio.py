"""Array files with the reference's API (vtem/utils/simpleio.py:10-149).

The reference keeps arrays in HDF5 files through PyTables: write_memory_to_file stores
root.real (and root.imag for complex input) as extendable arrays, write_array stores
root.array, and 'a' mode appends along the leading dimension.  Neither PyTables nor h5py
is part of this build, so the same nodes live in a zip of .npy members (numpy's .npz
container) under the file name the caller gives, with the node path as the member name
("real", "imag", "array").  Appending rewrites the file with the leading dimension grown.
Device tensors are copied to the host first, as the reference does with GPUArrays.
"""

from __future__ import annotations

import os
import zipfile

import numpy as np


def _host(A) -> np.ndarray:
    """ndarray of A (a numpy array or a torch tensor on any device)."""
    if isinstance(A, np.ndarray):
        return A
    if hasattr(A, "detach") and hasattr(A, "cpu"):
        return A.detach().cpu().numpy()
    raise TypeError("Write file error: unknown input")


def read_nodes(filename) -> dict:
    """Every node of an array file as {path: ndarray} (no pickled objects are loaded)."""
    with np.load(filename, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def write_nodes(filename, nodes: dict) -> None:
    """Write {path: ndarray} as the array file `filename` (no extension is added)."""
    tmp = f"{filename}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        np.savez(f, **nodes)
    os.replace(tmp, filename)


def _append(old: np.ndarray, new: np.ndarray, name: str) -> np.ndarray:
    if old.shape[1:] != new.shape[1:]:
        raise ValueError(f"append to {name}: the trailing dimensions {new.shape[1:]} differ from the "
                         f"file's {old.shape[1:]}")
    return np.concatenate([old, new.astype(old.dtype, copy=False)], axis=0)


def write_memory_to_file(A, filename, mode: str = "w", title: str = "test") -> None:
    """Store A as node "real" (and "imag" if A is complex); mode 'w' starts a new file, 'a'
    appends along the leading dimension (vtem/utils/simpleio.py:10-63)."""
    B = _host(A)
    if B.dtype not in (np.float32, np.float64, np.complex64, np.complex128, np.int32, np.int64):
        raise TypeError("Write file error: unknown input dtype")
    if B.ndim == 0:
        B = B.reshape(1)
    if np.iscomplexobj(B):
        nodes = {"real": np.ascontiguousarray(B.real), "imag": np.ascontiguousarray(B.imag)}
    else:
        nodes = {"real": np.ascontiguousarray(B)}
    if mode == "a" and os.path.exists(filename):
        old = read_nodes(filename)
        if set(old) != set(nodes):
            raise ValueError("append: the file and the input differ in being complex")
        nodes = {k: _append(old[k], v, k) for k, v in nodes.items()}
    elif mode not in ("w", "a"):
        raise ValueError("mode must be 'w' or 'a'")
    write_nodes(filename, nodes)


def write_array(A, filename, mode: str = "w", title: str = "test") -> None:
    """Store A (real or complex) as node "array" (vtem/utils/simpleio.py:65-125)."""
    B = _host(A)
    if B.dtype not in (np.float32, np.float64, np.complex64, np.complex128, np.int32, np.int64):
        raise TypeError("Write file error: unknown input dtype")
    if B.ndim == 0:
        B = B.reshape(1)
    B = np.ascontiguousarray(B)
    if mode == "a" and os.path.exists(filename):
        B = _append(read_nodes(filename)["array"], B, "array")
    elif mode not in ("w", "a"):
        raise ValueError("mode must be 'w' or 'a'")
    write_nodes(filename, {"array": B})


def read_file(filename) -> np.ndarray:
    """The array of a write_memory_to_file file: real, or real + 1j imag
    (vtem/utils/simpleio.py:127-149)."""
    nodes = read_nodes(filename)
    if set(nodes) == {"real"}:
        return nodes["real"]
    if set(nodes) == {"real", "imag"}:
        return nodes["real"] + np.array(1j).astype(np.complex64) * nodes["imag"]
    raise ValueError(f"{filename} is not an array file (nodes {sorted(nodes)})")


def read_array(filename) -> np.ndarray:
    """The array of a write_array file (vtem/utils/simpleio.py:151-162)."""
    return read_nodes(filename)["array"]


def is_array_file(filename) -> bool:
    """True if filename is a zip container of .npy members."""
    return os.path.isfile(filename) and zipfile.is_zipfile(filename)
