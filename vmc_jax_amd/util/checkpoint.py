"""Checkpoint / resume of a VMC run.

The reference stores the flat real parameter vector per time in HDF5
(output_manager.py:118-174) and re-thermalises on resume because walkers and
PRNG keys are not saved.  Here the same flat vector is written together with
the Metropolis walkers, the Philox seed and the sampler call counter, so a
resumed run draws exactly the proposals the uninterrupted run would have drawn.
Files are plain ``.npz`` archives (no pickled objects).
"""
import numpy as np
import torch

from .. import global_defs


def save(path, psi, sampler=None, time=0.0, extra=None):
    data = {"parameters": psi.get_parameters().cpu().numpy(), "time": np.float64(time)}
    if sampler is not None and hasattr(sampler, "states"):
        data["walkers"] = sampler.states.cpu().numpy()
        data["seed"] = np.uint64(sampler.key & 0xFFFFFFFFFFFFFFFF)
        data["call_index"] = np.int64(sampler.callIndex)
    if sampler is not None and hasattr(sampler, "lastNorm"):
        data["last_norm"] = np.float64(sampler.lastNorm)
    for k, v in (extra or {}).items():
        data["extra_" + k] = np.asarray(v)
    np.savez(path, **data)


def load(path, psi, sampler=None):
    """Restores the state written by ``save``; returns the saved time."""
    with np.load(path, allow_pickle=False) as z:
        psi.set_parameters(z["parameters"])
        if sampler is not None and "walkers" in z.files:
            w = torch.as_tensor(z["walkers"], device=global_defs.device()).contiguous()
            if tuple(w.shape) != tuple(sampler.states.shape):
                raise ValueError("walker shape %s does not match the sampler's %s"
                                 % (tuple(w.shape), tuple(sampler.states.shape)))
            sampler.states = w
            sampler.key = int(z["seed"])
            sampler.callIndex = int(z["call_index"])
        if sampler is not None and "last_norm" in z.files:
            sampler.lastNorm = float(z["last_norm"])
        return float(z["time"])
